// Writes the VTK files the Python test parses: Model3D / Model2D dumps and
// direct use of the rectilinear-grid writers (binary and ASCII, natural and
// XYZ orderings, a 2D vector dataset).  Usage: vtk_host <directory>.
#include "eikonalxx_amd.hpp"

#include <cstdio>
#include <string>
#include <vector>

int main(int argc, char **argv)
{
    using namespace EikonalXX;
    if (argc != 2) { std::printf("usage: vtk_host <dir>\n"); return 2; }
    const std::string dir = argv[1];

    Geometry3D g;
    g.setNumberOfGridPointsInX(4);
    g.setNumberOfGridPointsInY(3);
    g.setNumberOfGridPointsInZ(5);
    g.setGridSpacingInX(10.0);
    g.setGridSpacingInY(20.0);
    g.setGridSpacingInZ(5.0);
    g.setOriginInX(1.0);
    g.setOriginInY(2.0);
    g.setOriginInZ(3.0);
    std::vector<double> v(4*3*5);
    for (size_t i = 0; i < v.size(); ++i) { v[i] = 1000.0 + 7.0*i; }
    Model3D<double> m;
    m.initialize(g);
    m.setNodalVelocities((int) v.size(), v.data(), Ordering3D::Natural);
    // into a directory the writer creates
    m.writeVTK(dir + "/sub/model3d.vtk");

    IO::VTKRectilinearGrid3D w;
    w.open(dir + "/grid3d_ascii.vtk", g, "my title", false);
    std::vector<double> nodes(v.size());
    for (size_t i = 0; i < nodes.size(); ++i) { nodes[i] = 0.25*i; }
    w.writeNodalDataset("t field", nodes.data());
    std::vector<float> cellsXYZ(3*2*4);
    for (size_t i = 0; i < cellsXYZ.size(); ++i) { cellsXYZ[i] = 1.5f*i; }
    w.writeCellularDataset("cells", cellsXYZ.data(), Ordering3D::XYZ);
    w.close();

    IO::VTKRectilinearGrid3D b;
    b.open(dir + "/grid3d_binary.vtk", g, "", true);
    b.writeNodalDataset("", nodes.data(), Ordering3D::Natural);
    std::vector<double> vec(3*nodes.size());
    for (size_t i = 0; i < vec.size(); ++i) { vec[i] = -0.5*i; }
    b.writeNodalVectorDataset("grad", vec.data());
    b.close();

    Geometry2D g2;
    g2.setNumberOfGridPointsInX(5);
    g2.setNumberOfGridPointsInZ(4);
    g2.setGridSpacingInX(2.5);
    g2.setGridSpacingInZ(4.0);
    g2.setOriginInX(-3.0);
    g2.setOriginInZ(1.5);
    std::vector<double> v2(20);
    for (size_t i = 0; i < v2.size(); ++i) { v2[i] = 1500.0 + 11.0*i; }
    Model2D<double> m2;
    m2.initialize(g2);
    m2.setNodalVelocities((int) v2.size(), v2.data(), Ordering2D::Natural);
    m2.writeVTK(dir + "/model2d.vtk", "vp model");

    IO::VTKRectilinearGrid2D w2;
    w2.open(dir + "/grid2d.vtk", g2, "field 2d", true);
    std::vector<double> n2(20), vec2(40);
    for (size_t i = 0; i < n2.size(); ++i) { n2[i] = 3.0*i; }
    for (size_t i = 0; i < vec2.size(); ++i) { vec2[i] = 0.1*i; }
    w2.writeNodalDataset("times", n2.data(), Ordering2D::XZ);
    w2.writeNodalVectorDataset("times_gradient", vec2.data());
    w2.close();
    std::printf("ok\n");
    return 0;
}
