/* Prints the size and field offsets of the C ABI's structs, for the test
 * that compares them with the Python ctypes mirror. */
#include <stddef.h>
#include <stdio.h>
#include "eikonalxx_amd.h"

#define FIELD(T, f) printf("%s %s %zu\n", #T, #f, offsetof(T, f))

int main(void)
{
    printf("ekx_solver_options size %zu\n", sizeof(ekx_solver_options));
    FIELD(ekx_solver_options, number_of_sweeps);
    FIELD(ekx_solver_options, convergence_tolerance);
    FIELD(ekx_solver_options, precision);
    FIELD(ekx_solver_options, chains);
    FIELD(ekx_solver_options, chunk_planes);
    FIELD(ekx_solver_options, launch_order);
    printf("ekx_solve_stats size %zu\n", sizeof(ekx_solve_stats));
    FIELD(ekx_solve_stats, updates);
    FIELD(ekx_solve_stats, iterations_run);
    FIELD(ekx_solve_stats, tiled);
    FIELD(ekx_solve_stats, chunk_planes);
    FIELD(ekx_solve_stats, chains);
    FIELD(ekx_solve_stats, timed_kernel_ms);
    FIELD(ekx_solve_stats, sweep_ms_fp64);
    printf("ekx_geometry3d size %zu\n", sizeof(ekx_geometry3d));
    FIELD(ekx_geometry3d, dx);
    FIELD(ekx_geometry3d, z0);
    printf("ekx_problem2d size %zu\n", sizeof(ekx_problem2d));
    FIELD(ekx_problem2d, slowness);
    FIELD(ekx_problem2d, source);
    printf("ekx_layered_gaussian size %zu\n", sizeof(ekx_layered_gaussian));
    FIELD(ekx_layered_gaussian, layer_velocity);
    FIELD(ekx_layered_gaussian, seed);
    return 0;
}
