// GPU check of the C++ front end: the reference's own usage pattern
// (testing/solver3d.cpp:697-757, testing/solver2d.cpp:1195-1302) through
// include/eikonalxx_amd.hpp.  Writes the 3D field, then the 2D field and
// its two station times, as raw doubles to the file named by argv[1].
#include "eikonalxx_amd.hpp"

#include <cstdio>
#include <vector>

int main(int argc, char **argv)
{
    if (argc != 2) { return 2; }
    using namespace EikonalXX;
    // 3D: 9 x 10 x 11, h = 100, origin (5, 6, 7), v = 3000, radius 2
    int nx = 9, ny = 10, nz = 11;
    double dx = 100, x0 = 5, y0 = 6, z0 = 7;
    SolverOptions options;
    options.setNumberOfSweeps(1);
    options.setFactoredEikonalEquationSolverRadius(2);
    options.setAlgorithm(SolverAlgorithm::FastSweepingMethod);
    Geometry3D geometry;
    geometry.setNumberOfGridPointsInX(nx);
    geometry.setNumberOfGridPointsInY(ny);
    geometry.setNumberOfGridPointsInZ(nz);
    geometry.setGridSpacingInX(dx);
    geometry.setGridSpacingInY(dx);
    geometry.setGridSpacingInZ(dx);
    geometry.setOriginInX(x0);
    geometry.setOriginInY(y0);
    geometry.setOriginInZ(z0);
    std::vector<double> vConstant(nx*ny*nz, 3000);
    Model3D<double> vModel;
    vModel.initialize(geometry);
    vModel.setNodalVelocities((int) vConstant.size(), vConstant.data(),
                              Ordering3D::Natural);
    Source3D source;
    source.setGeometry(geometry);
    source.setLocationInX(x0 + dx*(nx/2) + dx/4);
    source.setLocationInY(y0 + dx*(ny/2) - dx/4);
    source.setLocationInZ(z0 + dx*(nz/4));
    Solver3D<double> solver;
    solver.initialize(options, geometry);
    solver.setVelocityModel(vModel);
    solver.setSource(source);
    solver.solve();
    std::vector<double> t3 = solver.getTravelTimeField();
    if (solver.getTravelTimeFieldPointer()[5] != t3[5]) { return 3; }

    // 2D: testing/solver2d.cpp:1195-1302
    int nx2 = 21, nz2 = 20;
    double dx2 = 100, dz2 = 100.05;
    Geometry2D g2;
    g2.setNumberOfGridPointsInX(nx2);
    g2.setNumberOfGridPointsInZ(nz2);
    g2.setGridSpacingInX(dx2);
    g2.setGridSpacingInZ(dz2);
    std::vector<double> v2(nx2*nz2, 5000);
    Model2D<double> m2;
    m2.initialize(g2);
    m2.setNodalVelocities((int) v2.size(), v2.data(), Ordering2D::Natural);
    double xs = dx2*(nx2/2) + dx2/4, zs = dz2*(nz2/2);
    SolverOptions o2;
    o2.setNumberOfSweeps(0);
    o2.setFactoredEikonalEquationSolverRadius(3);
    o2.setAlgorithm(SolverAlgorithm::FastSweepingMethod);
    std::vector<Station2D> stations(2);
    for (auto &s : stations) { s.setGeometry(g2); }
    stations[0].setLocationInX(xs);
    stations[0].setLocationInZ(zs);
    stations[1].setLocationInX(xs + 2*dx2);
    stations[1].setLocationInZ(zs + 3.1*dz2);
    Solver2D<double> s2;
    s2.initialize(o2, g2);
    s2.setVelocityModel(m2);
    s2.setSource(std::make_pair(xs, zs));
    s2.setStations(stations);
    s2.solve();
    std::vector<double> t2 = s2.getTravelTimeField();
    std::vector<double> ts = s2.getTravelTimesAtStations();

    FILE *f = std::fopen(argv[1], "wb");
    if (!f) { return 4; }
    std::fwrite(t3.data(), sizeof(double), t3.size(), f);
    std::fwrite(t2.data(), sizeof(double), t2.size(), f);
    std::fwrite(ts.data(), sizeof(double), ts.size(), f);
    std::fclose(f);
    std::printf("ok\n");
    return 0;
}
