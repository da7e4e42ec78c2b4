/* DeePAW-HALF C ABI v1, served by the MI355X (gfx950) backend.
 *
 * Names, enum values, struct layouts and argument meaning are those of the
 * reference's versioned ABI (reference: include/half.h:1-194,
 * implemented in src/half_c_api.F90:48-588), so existing C and Fortran hosts
 * (include/half_api.f90, the VASP adapter) link against this library
 * unchanged.  Each declaration names the reference entry point it replaces.
 *
 * Backend selection: HALF_BACKEND_CUDA (and AUTO) select the HIP backend of
 * this library; HALF_BACKEND_CPU returns HALF_ERROR_UNAVAILABLE because this
 * library carries no CPU numerics.  Capabilities report HALF_CAP_CUDA only.
 * Internal GPU failures return HALF_ERROR_INTERNAL instead of aborting.
 */
#ifndef DEEPAW_HALF_H
#define DEEPAW_HALF_H

#include <stdint.h>
#ifndef __cplusplus
#include <complex.h>
#define HALF_COMPLEX double _Complex
#else
#define HALF_COMPLEX double /* two doubles per element (re, im) */
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define HALF_ABI_VERSION 1
#define HALF_VERSION_MAJOR 0
#define HALF_VERSION_MINOR 6
#define HALF_VERSION_PATCH 0

/* reference: include/half.h:16-24 */
enum half_status {
  HALF_SUCCESS = 0,
  HALF_ERROR_INVALID_ARGUMENT = 1,
  HALF_ERROR_IO = 2,
  HALF_ERROR_INVALID_HANDLE = 3,
  HALF_ERROR_CAPACITY = 4,
  HALF_ERROR_UNAVAILABLE = 5,
  HALF_ERROR_INTERNAL = 6
};

/* reference: include/half.h:26-53 */
enum half_xc { HALF_XC_LDA = 1, HALF_XC_PBE = 2 };
enum half_backend { HALF_BACKEND_AUTO = 0, HALF_BACKEND_CPU = 1, HALF_BACKEND_CUDA = 2 };
enum half_solver {
  HALF_SOLVER_EVD = 1,
  HALF_SOLVER_EVJ = 2,
  HALF_SOLVER_VASP = 3, /* index-range dense solve of the lowest NBANDS */
  HALF_SOLVER_ACC = 4   /* matrix-free all-band solver, H.psi on the GPU */
};
enum half_capability {
  HALF_CAP_CPU = 1 << 0,
  HALF_CAP_CUDA = 1 << 1,
  HALF_CAP_MPI = 1 << 2,
  HALF_CAP_HDF5 = 1 << 3,
  HALF_CAP_SPGLIB = 1 << 4,
  HALF_CAP_ESCN_API = 1 << 5
};
enum half_escn_flag { HALF_ESCN_INCLUDE_UNCERTAINTY = 1 << 0 };
enum half_density_normalization { HALF_DENSITY_RAW = 0, HALF_DENSITY_NORMALIZE_VALENCE = 1 };

typedef int64_t half_handle;

/* reference: include/half.h:64-75 */
typedef struct half_request_geometry_v1 {
  uint32_t struct_size;
  uint32_t flags;
  int32_t nions;
  int32_t ntypes;
  int32_t grid[3];
  int32_t reserved_i32;
  double lattice[9];
  const int32_t *species;
  const double *positions_fractional;
  uint64_t reserved[8];
} half_request_geometry_v1;

/* reference: include/half.h:80-90 */
typedef struct half_escn_request_v1 {
  uint32_t struct_size;
  uint32_t flags;
  int32_t nions;
  int32_t grid[3];
  int32_t reserved_i32;
  double cell[9];
  const int32_t *atomic_numbers;
  const double *positions_cartesian;
  uint64_t reserved[8];
} half_escn_request_v1;

/* reference: include/half.h:96-110 */
typedef struct half_escn_result_v1 {
  uint32_t struct_size;
  uint32_t flags;
  int32_t grid[3];
  int32_t reserved_i32;
  int64_t capacity;
  float *density;
  float *nu;
  float *alpha;
  float *beta;
  float *risk;
  double elapsed_seconds;
  uint64_t reserved[8];
} half_escn_result_v1;

/* replaces src/half_c_api.F90:48-50 */
int half_get_abi_version(void);
/* replaces src/half_c_api.F90:52-54 */
const char *half_get_version_string(void);
/* replaces src/half_c_api.F90:56-76 */
int half_get_capabilities(void);

/* replace src/half_escn_client.c (half_escn_health / half_escn_predict);
 * the network density source is not part of this library: both return
 * HALF_ERROR_UNAVAILABLE. */
int half_escn_health(const char *base_url, int timeout_seconds, char *error, int error_capacity);
int half_escn_predict(const char *base_url, const half_escn_request_v1 *request,
                      half_escn_result_v1 *result, int timeout_seconds, char *error,
                      int error_capacity);

/* replaces src/half_c_api.F90:213-240 (AUTO backend + EVD) */
int half_create_from_files(const char *charge_path, const char *potential_path, double encut_eV,
                           int xc, int use_uspp_dij, half_handle *handle, char *error,
                           int error_capacity);
/* replaces src/half_c_api.F90:242-277 */
int half_create_from_files_backend(const char *charge_path, const char *potential_path,
                                   double encut_eV, int xc, int use_uspp_dij, int backend,
                                   int solver, half_handle *handle, char *error,
                                   int error_capacity);
/* replaces src/half_c_api.F90:78-211; validates its arguments, then returns
 * HALF_ERROR_UNAVAILABLE (no network density source). */
int half_create_from_escn(const char *base_url, const char *potential_path,
                          const half_request_geometry_v1 *geometry, double encut_eV, int xc,
                          int use_uspp_dij, int backend, int solver, int normalization,
                          int timeout_seconds, half_handle *handle, char *error,
                          int error_capacity);
/* replaces src/half_c_api.F90:279-287 */
int half_destroy(half_handle handle, char *error, int error_capacity);

/* replaces src/half_c_api.F90:289-335.  Same validation and messages; in
 * addition the geometry becomes the context's cell (the fix recorded in the
 * reference's docs/validation/HFO2_NWAVE_BOUNDARY_MISMATCH.md:43-50): the
 * request must have the density file's grid, one type per potential dataset
 * and the file's ion count per type (else HALF_ERROR_INVALID_ARGUMENT, the
 * context unchanged); its lattice and positions then replace the file's for
 * the basis, projectors, Veff and D of every later call, and the context's
 * device operator handles become invalid. */
int half_set_request_geometry(half_handle handle, const half_request_geometry_v1 *geometry,
                              char *error, int error_capacity);
/* replaces src/half_c_api.F90:337-386 */
int half_get_request_geometry(half_handle handle, int32_t *nions, int32_t *ntypes, int32_t grid[3],
                              double lattice[9], int64_t species_capacity, int32_t *species,
                              int64_t positions_capacity, double *positions_fractional,
                              char *error, int error_capacity);
/* replaces src/half_c_api.F90:388-401; reports the request geometry's
 * lattice and volume once one is set. */
int half_get_system_info(half_handle handle, int *nions, int *ntypes, int grid[3],
                         double lattice[9], double *volume_A3, char *error, int error_capacity);

/* replaces src/half_c_api.F90:403-417 */
int half_get_basis_size(half_handle handle, const double kpoint[3], int64_t *npw, char *error,
                        int error_capacity);
/* replaces src/half_c_api.F90:419-439 */
int half_get_basis(half_handle handle, const double kpoint[3], int64_t capacity, int32_t *gvec,
                   double *kinetic_eV, char *error, int error_capacity);

/* replaces src/half_c_api.F90:441-462; assembled on the GPU (the reference's
 * CUDA backend returned HALF_ERROR_UNAVAILABLE here). */
int half_assemble_hs(half_handle handle, const double kpoint[3], int64_t ld, HALF_COMPLEX *h,
                     HALF_COMPLEX *s, char *error, int error_capacity);
/* replaces src/half_c_api.F90:464-487; matrix-free FP64 H.psi / S.psi on
 * the GPU (the reference's CUDA backend returned HALF_ERROR_UNAVAILABLE). */
int half_apply_hs(half_handle handle, const double kpoint[3], int nstates, int64_t ld,
                  const HALF_COMPLEX *psi, HALF_COMPLEX *hpsi, HALF_COMPLEX *spsi, char *error,
                  int error_capacity);
/* replaces src/half_c_api.F90:489-519 */
int half_solve_kpoint(half_handle handle, const double kpoint[3], int nbands,
                      double *eigenvalues_eV, HALF_COMPLEX *eigenvectors, int64_t ld,
                      double *overlap_min, double *overlap_max, char *error, int error_capacity);
/* replaces src/half_c_api.F90:521-588.  The bases are compared before the
 * solve; a mismatch message keeps the reference's text and adds both sizes
 * and the first G vector found in one basis only. */
int half_solve_kpoint_mapped(half_handle handle, const double kpoint[3], int nbands,
                             int64_t host_npw, const int32_t *host_gvec, double *eigenvalues_eV,
                             HALF_COMPLEX *eigenvectors, int64_t ld, double *overlap_min,
                             double *overlap_max, char *error, int error_capacity);

#ifdef __cplusplus
}
#endif
#endif
