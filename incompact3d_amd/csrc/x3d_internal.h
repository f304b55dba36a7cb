// Internal declarations of the x3d HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/x3d.h"

namespace x3d {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define X3D_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::x3d::fail(X3D_ERR_HIP, std::string(#call) + ": " +              \
                                          hipGetErrorString(e_));              \
  } while (0)

#define X3D_TRY(call)                                                          \
  do {                                                                         \
    int r_ = (call);                                                           \
    if (r_ != X3D_OK) return r_;                                               \
  } while (0)

// ---------------------------------------------------------------- operators
enum LineOpKind { OP_DER1 = 0, OP_DER2 = 1, OP_DERVP = 2, OP_INTERVP = 3, OP_DERPV = 4, OP_INTERPV = 5 };

// Explicit-stencil scalars of one direction (module derivX/Y/Z, schemes.f90).
template <typename T>
struct Sten {
  // first derivative (first_derivative, schemes.f90:443-599)
  T afi, bfi, alfai, af1, bf1, cf1, af2, afm, afn, bfn, cfn;
  // second derivative (second_derivative, schemes.f90:602-856)
  T asi, bsi, csi, dsi, alsai;
  T as1, bs1, cs1, ds1, as2, as3, bs3, as4, bs4, cs4;
  T astt, bstt, cstt, ast, bst, asm_, asn, bsn, csn, dsn;
  // staggered derivative / interpolation (interpolation, schemes.f90:860-1066)
  T aci, bci, alcai, aici, bici, cici, dici, ailcai;
};

// One tridiagonal system (Thomas factors of prepare(), schemes.f90:413-439)
// plus, for periodic lines, the Sherman-Morrison vector r and the
// denominator 1 + r(1) - alfa*r(n) of derive.f90:55-57.  On the device the
// per-point factors are packed as pk[4*i..4*i+3] = {fs, fw, ff*fw, r}: the
// forward sweep y_i = rhs_i - fs_i y_{i-1}, the backward sweep
// x_i = fw_i y_i - (ff_i fw_i) x_{i+1}, and the correction x_i -= s r_i.
template <typename T>
struct TriDev {
  const T *pk = nullptr;
  T alfa = 0, sm_den = 1;
  int n = 0;
  int periodic = 0;
};

// Host FP64 description of a tridiagonal set, with device FP64/FP32 copies.
struct TriSet {
  std::string name;             // e.g. "x:ff" (ffx,fsx,fwx)
  std::vector<double> ff, fs, fw, sm;
  double alfa = 0, sm_den = 1;
  int periodic = 0;
  double *d64 = nullptr;        // packed {fs, fw, ff*fw, r} x n, FP64
  float *d32 = nullptr;         // the same, FP32
  TriDev<double> dev64() const;
  TriDev<float> dev32() const;
};

// Upload the packed device copies of one set (allocates d64/d32).
int upload_set(TriSet &s);

struct DirCoef {
  int n = 0, nm = 0, bc1 = 0, bcn = 0;
  double d = 0, d2 = 0, len = 0;
  Sten<double> s;
  // named arrays (reference names without the direction letter) for
  // x3d_get_coefficients, and tridiagonal sets for the operators
  std::vector<std::pair<std::string, std::vector<double>>> arrays;
  std::vector<std::pair<std::string, double>> scalars;
  int set_d1_odd = -1, set_d1_even = -1, set_d2_odd = -1, set_d2_even = -1;
  int set_c6 = -1, set_ci6p = -1, set_ci6 = -1, set_cip6 = -1, set_ciip6 = -1;
};

struct Context {
  x3d_params p{};
  int nxm = 0, nym = 0, nzm = 0;
  double xnu = 0;
  DirCoef dir[3];
  std::vector<TriSet> sets;
  hipStream_t stream = nullptr;
  int device = 0;
  // Poisson state (fft.hip)
  void *poisson = nullptr;
  // scratch for host-memory entries
  void *host_stage = nullptr;
  size_t host_stage_bytes = 0;
};

Context *ctx();                 // nullptr if x3d_init not called
int require_ctx(Context **c);

// schemes.cpp
int build_schemes(Context &c);
int upload_sets(Context &c);
void free_sets(Context &c);
const TriSet *match_set(const Context &c, int dir, int n, const double *ff, const double *fs,
                        const double *fw);
const TriSet *match_set32(const Context &c, int dir, int n, const float *ff, const float *fs,
                          const float *fw);

// ---------------------------------------------------------------- line ops
template <typename T>
struct LineCall {
  int dir = 0;                  // 0 x, 1 y, 2 z
  int op = OP_DER1;
  int nx = 0, ny = 0, nz = 0;   // extents of the INPUT array
  int n_out = 0;                // output length along dir (other extents unchanged)
  const T *u = nullptr;         // input
  const T *u2 = nullptr;        // optional: input is u*u2 (fused product)
  T *t = nullptr;               // output
  TriDev<T> tri;
  Sten<T> s;
  int bc1 = 0, bcn = 0, npaire = 0;
  const T *pp = nullptr;        // optional post-multiplier along dir (stretching)
  // epilogue: t = ac*acc + rc*(mul*result)   (acc/mul optional, may alias t)
  const T *acc = nullptr;
  const T *mul = nullptr;
  T ac = T(1), rc = T(1);
};

template <typename T>
int launch_line(const LineCall<T> &c, hipStream_t st);

template <typename T>
Sten<T> cast_sten(const Sten<double> &s);

// Build a LineCall for operator `op` along `dir` using the library's own
// coefficient sets (the sets the reference callers pass).
template <typename T>
int make_line_call(const Context &c, int dir, int op, int npaire, LineCall<T> *out);

// ---------------------------------------------------------------- fields
template <typename T>
int launch_copy(const T *src, T *dst, size_t n, hipStream_t st);

// ---------------------------------------------------------------- poisson
int poisson_setup(Context &c);
void poisson_free(Context &c);
template <typename T>
int poisson_solve(Context &c, T *rhs, hipStream_t st);

// ---------------------------------------------------------------- fused sweeps
// momentum.hip: the transeq right-hand side of component comp along dir
// (out = acc + ...); the x sweep also applies AB3 (nterms > 0) into u_new.
bool momentum_fusable(const Context &c);
template <typename T>
int launch_momentum(const Context &c, int dir, int comp, const T *const u[3], const T *acc, T *out,
                    const T *f2, const T *f3, const double *u_old, double *u_new, float *shadow, double ca,
                    double cb, double cc, int nterms, int nx, int ny, int nz, hipStream_t st);

// projection.hip: divergence (x, y, z) and gradient + cor_vel (z, y, x)
enum { SW_DIV_X = 0, SW_DIV_Y, SW_DIV_Z, SW_GRAD_Z, SW_GRAD_Y, SW_GRAD_X };
bool projection_fusable(const Context &c);
template <typename T>
int launch_projection(const Context &c, int sw, const T *const in[3], T *const out[3], double *const u64[3],
                      float *const u32[3], int nx, int ny, int nz, hipStream_t st);

}  // namespace x3d
