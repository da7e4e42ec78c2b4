// Shared types of the MI355X PAW-operator library: status codes, physical
// constants, host-side crystal / dataset / basis records, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace half {

// Status codes of the versioned C ABI (reference: include/half.h:16-24).
enum Status : int {
  kSuccess = 0,
  kInvalidArgument = 1,
  kIo = 2,
  kInvalidHandle = 3,
  kCapacity = 4,
  kUnavailable = 5,
  kInternal = 6
};

// Every failure inside the library is an Error carrying an ABI status; the
// C entry points translate it into a return code and a message.  Nothing in
// the library aborts the host process.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HALF_HIP(call)                                                             \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::half::Error(::half::kInternal, std::string("HIP: ") +                 \
                                                hipGetErrorString(e_) + " in " #call); \
  } while (0)

// Physical constants in the reference's convention
// (reference: src/half_constants.F90:6-14).
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kAutoa = 0.529177249;
constexpr double kRytoev = 13.605826;
constexpr double kHatoev = 2.0 * kRytoev;
constexpr double kFelect = 2.0 * kRytoev * kAutoa;
constexpr double kEdeps = 4.0 * kPi * kFelect;
constexpr double kHsqdtm = kRytoev * kAutoa * kAutoa;

struct Crystal {
  std::string system;            // first line of the density file
  double lattice[3][3] = {};     // rows are lattice vectors, Angstrom
  double reciprocal[3][3] = {};  // rows b_i = 2 pi (A^-1)^T
  double volume = 0;
  std::vector<std::string> species;
  std::vector<int> counts;
  std::vector<double> positions;  // fractional, nions x 3, grouped by species
  int nions() const {
    int n = 0;
    for (int c : counts) n += c;
    return n;
  }
  void update_geometry();  // reference: src/half_types.F90:59-63
};

// One POTCAR dataset, restricted to the fields the operator path consumes
// (reference: src/half_types.F90:41-56).
struct Dataset {
  std::string element;
  double zval = 0, eatom = 0;
  double psp_gmax = 0, pspnl_gmax = 0;
  bool has_core = false;
  std::vector<double> psp_local, pspcor;  // 1000-point tables
  std::vector<int> lps;                   // l of each channel
  std::vector<double> dion, qpaw;         // channels x channels, row-major
  std::vector<double> pspnl;              // channels x 100, channel-major
  // PAW radial sets: augmentation-sphere radius, radial grid and the
  // r-weighted pseudo / all-electron partial waves (channel-major, nmax each)
  double paw_rmax = 0;
  std::vector<double> rgrid, wps, wae;
  int channels() const { return static_cast<int>(lps.size()); }
  int nlm() const {
    int n = 0;
    for (int l : lps) n += 2 * l + 1;
    return n;
  }
};

struct ChargeGrid {
  int shape[3] = {0, 0, 0};
  std::vector<double> values;  // as stored in the file: x fastest
  int64_t size() const { return int64_t(shape[0]) * shape[1] * shape[2]; }
};

// Plane-wave basis at one k point (reference: src/half_types.F90:31-39).
struct Basis {
  int shape[3] = {0, 0, 0};
  double encut = 0;
  double kpoint[3] = {0, 0, 0};
  std::vector<int32_t> gvec;      // npw x 3
  std::vector<double> q;          // npw x 3, Cartesian k+G
  std::vector<double> kinetic;    // npw
  std::vector<int32_t> fft_index; // npw, zero-based C-order box index
  int64_t npw() const { return static_cast<int64_t>(kinetic.size()); }
};

// Readers (reference: src/half_chgcar.F90:12-98, src/half_potcar.F90:24-205).
void read_chgcar(const std::string &path, Crystal &crystal, ChargeGrid &charge);
std::vector<Dataset> read_potcar(const std::string &path);
void validate_datasets(const std::vector<Dataset> &sets, const Crystal &crystal);

// reference: src/half_basis.F90:9-55
Basis build_basis(const Crystal &crystal, const int shape[3], double encut,
                  const double kpoint[3]);

// Uniform-grid cubic spline second derivatives with the reference's end
// conditions (reference: src/half_paw.F90:78-99).
std::vector<double> spline_second(const std::vector<double> &y, double xmax);

// lm expansion of DION/QPAW: D(ilm,jlm) = DION(ich,jch) iff l and m match
// (reference: src/half_paw.F90:39-51).  Row-major nlm x nlm.
void lm_expand(const Dataset &d, std::vector<double> &dlm, std::vector<double> &qlm);

// RAII device allocation.
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  ~DeviceArray() { release(); }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  DeviceArray(DeviceArray &&o) noexcept : ptr_(o.ptr_), n_(o.n_) {
    o.ptr_ = nullptr;
    o.n_ = 0;
  }
  DeviceArray &operator=(DeviceArray &&o) noexcept {
    if (this != &o) {
      release();
      ptr_ = o.ptr_;
      n_ = o.n_;
      o.ptr_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) HALF_HIP(hipMalloc(reinterpret_cast<void **>(&ptr_), n * sizeof(T)));
    n_ = n;
  }
  void ensure(size_t n) {
    if (n > n_) resize(n);
  }
  void upload(const T *host, size_t n, hipStream_t s) {
    ensure(n);
    if (n) HALF_HIP(hipMemcpyAsync(ptr_, host, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
  T *get() const { return ptr_; }
  size_t size() const { return n_; }
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    n_ = 0;
  }

 private:
  T *ptr_ = nullptr;
  size_t n_ = 0;
};

}  // namespace half
