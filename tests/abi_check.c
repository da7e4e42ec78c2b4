/* C ABI boundary, checked from plain C with no GPU needed.
 *
 * Sections:
 *   1. version and capability words;
 *   2. every entry point that takes a context or operator handle rejects a
 *      stale one with HALF_ERROR_INVALID_HANDLE and says so in the error
 *      buffer (the handle is checked before any pointer argument is read);
 *   3. error buffers: NUL-terminated, truncated to capacity - 1 characters,
 *      left alone when NULL or of capacity 0;
 *   4. context creation: selector and path validation, missing inputs are an
 *      IO error, the CPU backend is unavailable, the handle output is zeroed
 *      on failure; the network density source is unavailable.
 * argv[1], argv[2]: an existing density and potential file (section 4).
 *
 * The reference's own tests/test_c_api.c is compiled unchanged against this
 * library by oracle/Makefile and run by tests/test_abi.py beside this one. */
#include <stdio.h>
#include <string.h>

#include "half.h"
#include "half_amd.h"

static int failures = 0;

static void expect_status(const char *what, int got, int want) {
  if (got != want) {
    fprintf(stderr, "%s: status %d, expected %d\n", what, got, want);
    ++failures;
  }
}

static void expect_true(const char *what, int ok) {
  if (!ok) {
    fprintf(stderr, "%s: failed\n", what);
    ++failures;
  }
}

static char err[256];

/* a call on a stale handle: status and message */
#define STALE(call)                                                                      \
  do {                                                                                   \
    memset(err, 0, sizeof err);                                                          \
    expect_status(#call, (call), HALF_ERROR_INVALID_HANDLE);                             \
    expect_true("message of " #call, strstr(err, "invalid") != NULL && strstr(err, "handle") != NULL); \
  } while (0)

int main(int argc, char **argv) {
  const half_handle stale = 12345;          /* never issued */
  const half_amd_operator stale_op = 777;  /* never issued */
  const double k0[3] = {0.0, 0.0, 0.0};
  const int32_t g0[3] = {0, 0, 0};
  double d[16] = {0};
  int32_t i32[16] = {0};
  int64_t i64[4] = {0};
  int ints[8] = {0};
  double _Complex z[4] = {0};
  const double _Complex *zs[1] = {z};
  half_amd_solve_stats stats;
  half_amd_harris_terms terms;
  half_amd_operator op = 0;
  half_handle handle = 99;

  /* 1 */
  expect_status("ABI version", half_get_abi_version(), HALF_ABI_VERSION);
  expect_true("version string 0.6.0", strcmp(half_get_version_string(), "0.6.0") == 0);
  expect_true("only known capability bits",
              (half_get_capabilities() & ~(HALF_CAP_CPU | HALF_CAP_CUDA | HALF_CAP_MPI | HALF_CAP_HDF5 |
                                           HALF_CAP_SPGLIB | HALF_CAP_ESCN_API)) == 0);
  expect_true("GPU backend advertised", (half_get_capabilities() & HALF_CAP_CUDA) != 0);

  /* 2: context handles (half.h), then the library's own entry points */
  STALE(half_destroy(stale, err, sizeof err));
  STALE(half_set_request_geometry(stale, NULL, err, sizeof err));
  STALE(half_get_request_geometry(stale, &i32[0], &i32[1], &i32[2], d, 4, &i32[5], 4, d, err, sizeof err));
  STALE(half_get_system_info(stale, &ints[0], &ints[1], &ints[2], d, &d[9], err, sizeof err));
  STALE(half_get_basis_size(stale, k0, &i64[0], err, sizeof err));
  STALE(half_get_basis(stale, k0, 1, i32, d, err, sizeof err));
  STALE(half_assemble_hs(stale, k0, 1, z, z, err, sizeof err));
  STALE(half_apply_hs(stale, k0, 1, 1, z, z, z, err, sizeof err));
  STALE(half_solve_kpoint(stale, k0, 1, d, z, 1, &d[1], &d[2], err, sizeof err));
  STALE(half_solve_kpoint_mapped(stale, k0, 1, 1, g0, d, z, 1, &d[1], &d[2], err, sizeof err));
  STALE(half_amd_operator_create(stale, k0, 64, &op, &i64[0], &i32[0], err, sizeof err));
  STALE(half_amd_solve_acc(stale, k0, 1, 0, 1e-4, 0.0, 40, 0, 0, d, z, 1, &stats, err, sizeof err));
  STALE(half_amd_solve_history(stale, 4, d, &i32[0], err, sizeof err));
  STALE(half_amd_solve_kmesh(stale, 1, k0, 1, 0, 1e-4, 0.0, 40, 0, 1, d, &stats, err, sizeof err));
  STALE(half_amd_force_set(stale, 1, k0, d, 1, d, d, zs, i64, d, d, d, d, err, sizeof err));
  STALE(half_amd_write_vaspwave(stale, "unused.h5", 1, k0, 1, d, d, 0.0, zs, i64, err, sizeof err));
  STALE(half_amd_set_solver(stale, HALF_SOLVER_EVD, err, sizeof err));
  STALE(half_amd_energy_terms(stale, &d[0], &d[1], &d[2], err, sizeof err));
  STALE(half_amd_harris_energy(stale, 1, d, 1, d, 0.0, 0, d, &terms, err, sizeof err));
  STALE(half_amd_get_veff(stale, 1, d, err, sizeof err));
  STALE(half_amd_nonlocal_forces(stale, k0, 1, d, d, 1.0, z, 1, d, err, sizeof err));
  STALE(half_amd_set_projector_positions(stale, 1, d, err, sizeof err));
  /* operator handles */
  STALE(half_amd_operator_apply(stale_op, 0, 1, z, 1, z, z, 1, 0, NULL, err, sizeof err));
  STALE(half_amd_operator_apply_noalloc(stale_op, 1, z, 1, z, z, 1, NULL, err, sizeof err));
  STALE(half_amd_operator_destroy(stale_op, err, sizeof err));

  /* 3 */
  {
    char tiny[8];
    memset(tiny, 'x', sizeof tiny);
    expect_status("truncated message", half_destroy(stale, tiny, sizeof tiny), HALF_ERROR_INVALID_HANDLE);
    expect_true("truncated to capacity - 1 plus NUL", strlen(tiny) == sizeof tiny - 1 && strncmp(tiny, "invalid", 7) == 0);
    memset(tiny, 'x', sizeof tiny);
    expect_status("capacity 0", half_destroy(stale, tiny, 0), HALF_ERROR_INVALID_HANDLE);
    expect_true("capacity 0 leaves the buffer alone", tiny[0] == 'x' && tiny[7] == 'x');
    expect_status("NULL buffer", half_destroy(stale, NULL, 64), HALF_ERROR_INVALID_HANDLE);
  }

  /* 4 */
  handle = 99;
  expect_status("empty paths", half_create_from_files_backend("", "", 400.0, HALF_XC_PBE, 0, HALF_BACKEND_AUTO,
                                                               HALF_SOLVER_EVD, &handle, err, sizeof err),
                HALF_ERROR_INVALID_ARGUMENT);
  expect_true("handle zeroed on failure", handle == 0);
  expect_status("backend selector 99", half_create_from_files_backend("a", "b", 400.0, HALF_XC_PBE, 1, 99,
                                                                       HALF_SOLVER_EVD, &handle, err, sizeof err),
                HALF_ERROR_INVALID_ARGUMENT);
  expect_status("XC selector 3", half_create_from_files_backend("a", "b", 400.0, 3, 0, HALF_BACKEND_AUTO,
                                                                 HALF_SOLVER_EVD, &handle, err, sizeof err),
                HALF_ERROR_INVALID_ARGUMENT);
  expect_status("solver selector 5", half_create_from_files_backend("a", "b", 400.0, HALF_XC_PBE, 0,
                                                                     HALF_BACKEND_AUTO, 5, &handle, err, sizeof err),
                HALF_ERROR_INVALID_ARGUMENT);
  handle = 99;
  expect_status("missing inputs", half_create_from_files_backend("/nonexistent/CHGCAR", "/nonexistent/POTCAR", 400.0,
                                                                  HALF_XC_PBE, 0, HALF_BACKEND_AUTO, HALF_SOLVER_ACC,
                                                                  &handle, err, sizeof err),
                HALF_ERROR_IO);
  expect_true("handle zeroed on IO failure", handle == 0);
  if (argc > 2) {
    expect_status("non-positive ENCUT", half_create_from_files_backend(argv[1], argv[2], -1.0, HALF_XC_PBE, 0,
                                                                        HALF_BACKEND_AUTO, HALF_SOLVER_EVD, &handle,
                                                                        err, sizeof err),
                  HALF_ERROR_INVALID_ARGUMENT);
    expect_status("CPU backend", half_create_from_files_backend(argv[1], argv[2], 400.0, HALF_XC_PBE, 0,
                                                                 HALF_BACKEND_CPU, HALF_SOLVER_EVD, &handle, err,
                                                                 sizeof err),
                  HALF_ERROR_UNAVAILABLE);
    expect_true("CPU backend message", strstr(err, "CPU") != NULL);
  }
  expect_status("network density source", half_escn_health("http://127.0.0.1:1", 1, err, sizeof err),
                HALF_ERROR_UNAVAILABLE);

  if (failures) {
    fprintf(stderr, "%d boundary checks failed\n", failures);
    return 1;
  }
  puts("C ABI boundary: ok");
  return 0;
}
