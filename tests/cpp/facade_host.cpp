// Host-only checks of the C++ front end (include/eikonalxx_amd.hpp): the
// parts of the reference API that never touch a GPU.  Prints "ok" and
// exits 0 when every check holds; prints the failing check otherwise.
#include "eikonalxx_amd.hpp"

#include <cmath>
#include <cstdio>
#include <stdexcept>

#define CHECK(cond)                                                        \
    do                                                                     \
    {                                                                      \
        if (!(cond))                                                       \
        {                                                                  \
            std::printf("FAILED %s:%d %s\n", __FILE__, __LINE__, #cond);   \
            return 1;                                                      \
        }                                                                  \
    } while (0)

template <class E, class F> bool throws(F f)
{
    try { f(); } catch (const E &) { return true; } catch (...) {}
    return false;
}

int main()
{
    using namespace EikonalXX;
    // SolverOptions defaults (src/solverOptions.cpp:5-21) and setters
    // (testing/solverOptions.cpp).
    SolverOptions options;
    CHECK(options.getNumberOfSweeps() == 5);
    CHECK(options.getFactoredEikonalEquationSolverRadius() == 3);
    CHECK(options.getAlgorithm() == SolverAlgorithm::LevelSetMethod);
    CHECK(std::abs(options.getConvergenceTolerance() - 1.e-5) < 1.e-15);
    options.setConvergenceTolerance(1.e-3);
    options.setNumberOfSweeps(12);
    options.setFactoredEikonalEquationSolverRadius(6);
    options.setAlgorithm(SolverAlgorithm::FastSweepingMethod);
    options.setVerbosity(Verbosity::Info);
    SolverOptions copy(options);
    CHECK(copy.getNumberOfSweeps() == 12);
    CHECK(copy.getFactoredEikonalEquationSolverRadius() == 6);
    CHECK(std::abs(copy.getConvergenceTolerance() - 1.e-3) < 1.e-10);
    CHECK(copy.getAlgorithm() == SolverAlgorithm::FastSweepingMethod);
    CHECK(copy.getVerbosity() == Verbosity::Info);

    // Geometry validation (src/geometry3d.cpp:73-261)
    Geometry3D g;
    CHECK(throws<std::invalid_argument>([&] { g.setNumberOfGridPointsInX(2); }));
    CHECK(throws<std::invalid_argument>([&] { g.setGridSpacingInY(0.0); }));
    CHECK(throws<std::runtime_error>([&] { (void) g.getGridSpacingInZ(); }));
    g.setNumberOfGridPointsInX(21);
    g.setNumberOfGridPointsInY(18);
    g.setNumberOfGridPointsInZ(20);
    g.setGridSpacingInX(100);
    g.setGridSpacingInY(200);
    g.setGridSpacingInZ(100.05);
    g.setOriginInX(100);
    g.setOriginInY(-10);
    g.setOriginInZ(200);
    CHECK(g.getNumberOfGridPoints() == 21*18*20);
    CHECK(g.getNumberOfCells() == 20*17*19);

    // Source3D (testing/source.cpp:62-126)
    Source3D source;
    source.setGeometry(g);
    double xs = 100 + 100*(21/2) + 100/4.0;
    double ys = -10 + 200*(18/2) - 200/4.0;
    double zs = 200 + 100.05*(20/2);
    source.setLocationInX(xs);
    source.setLocationInY(ys);
    source.setLocationInZ(zs);
    CHECK(source.getCellInX() == 10);
    CHECK(source.getCellInY() == 8);
    CHECK(source.getCellInZ() == 10);
    CHECK(source.getCell() == 10*20*17 + 8*20 + 10);
    CHECK(std::abs(source.getOffsetInX() - (xs - 100)) < 1.e-14);
    source.setZToFreeSurface();
    CHECK(source.getCellInZ() == 0);
    CHECK(throws<std::invalid_argument>([&] { source.setLocationInX(99); }));

    // Model3D (testing/model.cpp:174-266): nodal -> cell slowness
    Geometry3D mg;
    mg.setNumberOfGridPointsInX(5);
    mg.setNumberOfGridPointsInY(4);
    mg.setNumberOfGridPointsInZ(6);
    mg.setGridSpacingInX(1);
    mg.setGridSpacingInY(1);
    mg.setGridSpacingInZ(1);
    Model3D<double> model;
    model.initialize(mg);
    std::vector<double> v(5*4*6), vt(5*4*6);
    for (int iz = 0; iz < 6; ++iz)
        for (int iy = 0; iy < 4; ++iy)
            for (int ix = 0; ix < 5; ++ix)
            {
                v[(iz*4 + iy)*5 + ix] = (iz*4 + iy)*5 + ix + 1;
                vt[(ix*4 + iy)*6 + iz] = (iz*4 + iy)*5 + ix + 1;
            }
    model.setNodalVelocities((int) v.size(), v.data(), Ordering3D::Natural);
    double avg = 0;
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i)
                avg += v[((2 + k)*4 + (1 + j))*5 + (3 + i)];
    avg /= 8;
    CHECK(std::abs(model.getSlowness(3, 1, 2) - 1/avg) < 1.e-12);
    model.setNodalVelocities((int) vt.size(), vt.data(), Ordering3D::XYZ);
    CHECK(std::abs(model.getSlowness(3, 1, 2) - 1/avg) < 1.e-12);
    v[7] = 0;
    CHECK(throws<std::invalid_argument>(
        [&] { model.setNodalVelocities((int) v.size(), v.data(),
                                       Ordering3D::Natural); }));
    CHECK(throws<std::invalid_argument>(
        [&] { model.setNodalVelocities(3, v.data(), Ordering3D::Natural); }));

    // Solver order violations (src/solver3d.cpp:207-404) -- no GPU needed
    Solver3D<double> solver;
    CHECK(!solver.isInitialized());
    CHECK(throws<std::runtime_error>([&] { solver.solve(); }));
    CHECK(throws<std::runtime_error>(
        [&] { solver.setSource(std::make_tuple(1.0, 1.0, 1.0)); }));
    Geometry3D incomplete;
    incomplete.setNumberOfGridPointsInX(5);
    CHECK(throws<std::invalid_argument>(
        [&] { solver.initialize(options, incomplete); }));
    Solver2D<double> solver2;
    CHECK(throws<std::runtime_error>([&] { solver2.solve(); }));
    std::printf("ok\n");
    return 0;
}
