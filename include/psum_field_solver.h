/*
 * psum_field_solver.h -- PSuM's field-solver plugin ABI, served by the
 * MI355X library.
 *
 * The typedefs, the func_table layout and the two exported registry
 * functions are the reference's binary contract
 * (src/field_solver/register/interface.h:4-28, registry.cpp:6-23), so a
 * PSuM Poisson_solver_{2,3}d frontend (solver_backend, backend.hpp:11-107)
 * can select the HIP multigrid by name with no source change:
 *
 *   "hip_multigrid_gpu"     FP64 Galerkin V-cycles (replaces
 *                           sycl_multigrid_gpu.cpp:22-230,
 *                           cuda_multigrid_gpu.cu:26-337)
 *   "hip_multigrid32f_gpu"  FP32 V-cycles in an FP64 residual refinement
 *                           loop (replaces sycl_multigrid32f_gpu.cpp:23-266)
 *
 * Contract kept from the reference:
 *   set_options  whitespace "key value" pairs; I J K depth pre_relax
 *                post_relax max_cycles correction_interval and the
 *                num_conv_mask* keys (accepted, unused: the HIP operator is
 *                matrix-free); a missing K means K = 1; unknown keys ignored.
 *                Extensions: "tol T" stops when ||b-Ax||/||b|| <= T (checked
 *                after each cycle, at most max_cycles); "coarse_fp64 0|1".
 *   set_matrix   host COO arrays, copied; call after set_options.
 *   set_source_addback / set_source_replace   host arrays, copied; applied
 *                to b in place on every solve, addback first.
 *   solve(b, x)  device pointers; x is the warm start and the result.
 * Errors: C++ exceptions (std::runtime_error) propagate out of these
 * functions exactly as in the reference backends.
 */
#ifndef PSUM_FIELD_SOLVER_H
#define PSUM_FIELD_SOLVER_H

#ifdef __cplusplus
extern "C" {
#endif

typedef void* psum_field_solver_handle;
typedef psum_field_solver_handle (*solver_init_func)();
typedef void (*solver_set_matrix_func)(psum_field_solver_handle, unsigned long long,
                                       unsigned long long, unsigned long long,
                                       unsigned long long*, unsigned long long*, double*);
typedef void (*solver_solve_func)(psum_field_solver_handle, double*, double*);
typedef void (*solver_set_options_func)(psum_field_solver_handle, const char*);
typedef void (*solver_set_source_replace_func)(psum_field_solver_handle, unsigned long long,
                                               unsigned long long*, double*);
typedef void (*solver_set_source_addback_func)(psum_field_solver_handle, unsigned long long,
                                               unsigned long long*, double*);
typedef void (*solver_free_func)(psum_field_solver_handle);

struct func_table {
    solver_init_func init;
    solver_set_matrix_func set_matrix;
    solver_solve_func solve;
    solver_set_options_func set_options;
    solver_set_source_replace_func set_source_replace;
    solver_set_source_addback_func set_source_addback;
    solver_free_func free;
};

/* registry (interface.h:26-27) */
__attribute__((visibility("default"))) void register_solver(const char* name, struct func_table table);
__attribute__((visibility("default"))) struct func_table get_solver_funcs(const char* name);

/* Non-throwing wrappers over the same backends for C callers and ctypes:
 * return PSUM_OK (0) or a psum_status; the message is in psum_hip_last_error().
 * `kind` is a registered backend name. */
__attribute__((visibility("default"))) void* psum_solver_create(const char* kind);
__attribute__((visibility("default"))) int psum_solver_set_options(void* h, const char* options);
__attribute__((visibility("default"))) int psum_solver_set_matrix(void* h, unsigned long long n_row,
        unsigned long long n_col, unsigned long long nnz, unsigned long long* rows,
        unsigned long long* cols, double* vals);
__attribute__((visibility("default"))) int psum_solver_set_source_replace(void* h, unsigned long long n,
        unsigned long long* idxs, double* vals);
__attribute__((visibility("default"))) int psum_solver_set_source_addback(void* h, unsigned long long n,
        unsigned long long* idxs, double* vals);
__attribute__((visibility("default"))) int psum_solver_solve(void* h, double* b_dev, double* x_dev);
/* cycles run by the last solve and its final ||b-Ax||/||b|| (NaN if not computed) */
__attribute__((visibility("default"))) int psum_solver_stats(void* h, int* cycles, double* rel_res);
__attribute__((visibility("default"))) void psum_solver_free(void* h);

#ifdef __cplusplus
}
#endif

#endif
