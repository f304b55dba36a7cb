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
  T alfa = 0, sm_den = 1, rden = 1;   // rden = 1 / sm_den
  int n = 0;
  int periodic = 0;
  int kdepth = 0;               // segment carry-chain depth (0: full chain)
};

// Host FP64 description of a tridiagonal set, with device FP64/FP32 copies.
struct TriSet {
  std::string name;             // e.g. "x:ff" (ffx,fsx,fwx)
  std::vector<double> ff, fs, fw, sm;
  double alfa = 0, sm_den = 1;
  int periodic = 0;
  double *d64 = nullptr;        // packed {fs, fw, ff*fw, r} x n, FP64
  float *d32 = nullptr;         // the same, FP32
  int kd64 = 0, kd32 = 0;       // chain depths of the FP64 and FP32 segmented solvers
  TriDev<double> dev64() const;
  TriDev<float> dev32() const;
};

// Segment lengths of the segmented y/z solvers in FP64 and FP32.
// (8-point FP32 segments were measured slower: 72 us against 59 us for the
// fused FP32 momentum sweep at 256^3, from the doubled carry chain.)
constexpr int SEG_L64 = 8, SEG_L32 = 16;

// Upload the packed device copies of one set (allocates d64/d32).
int upload_set(TriSet &s);
// Same coefficients (ff, fs, fw, sm, alfa, denominator).
bool same_set(const TriSet &a, const TriSet &b);
// Segment carry-chain depth for L-point segments at tolerance tol.
int chain_depth(const TriSet &s, int L, double tol);

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
  // stretched y mesh (istret != 0): the metrics of stretching.f90 on the
  // velocity (ppy) and pressure (ppyi) grids, host FP64 and device copies
  double alpha = 0;
  std::vector<double> yp, ypi, ppy, ppyi, pp2y, pp4y;
  double *ppy64 = nullptr, *ppyi64 = nullptr;
  float *ppy32 = nullptr, *ppyi32 = nullptr;
  // the y viscous term of transeq.f90:337-370 on a stretched mesh,
  // pp2y D2 - pp4y (ppy D1): device copies of pp2y and of pp4y*ppy
  double *pp2y64 = nullptr, *pp4yy64 = nullptr;
  float *pp2y32 = nullptr, *pp4yy32 = nullptr;
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
int compute_stretching(Context &c); // istret != 0: the metrics on the host
int build_stretching(Context &c);   // ... and their device copies
void free_stretching(Context &c);
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
// the solve on a caller-provided spectral work array of poisson_work_bytes
template <typename T>
size_t poisson_work_bytes(const Context &c);
template <typename T>
int poisson_run(const Context &c, T *rhs, void *work, hipStream_t st);
// device address of the file-scope scratch array on `device` (taken at
// x3d_init), nullptr before; POISSON_SCRATCH_BYTES bytes
// decomp_2d_fft_3d on the pressure grid: real (nx,ny,nz) <-> complex
// (nx, ny, nz/2+1), unnormalised
template <typename T>
int fft3d(Context &c, bool forward, const void *in, void *out, hipStream_t st);
constexpr size_t POISSON_SCRATCH_BYTES = (size_t)160 << 20;
void *poisson_scratch(int device);

// ---------------------------------------------------------------- fused sweeps
// momentum.hip: the transeq right-hand side of component comp along dir
// (out = acc + ...); the x sweep also applies AB3 (nterms > 0): FP64 compute
// writes the new velocity to u_new; FP32 compute writes the FP32 increment
// to delta and the new FP32 velocity to shadow, and the FP64 velocity takes
// the increment in the gradient correction.
bool momentum_fusable(const Context &c);
template <typename T>
int launch_momentum(const Context &c, int dir, int comp, const T *const u[3], const T *acc, T *out,
                    const T *f2, const T *f3, double *u_new, float *delta, float *shadow, double ca,
                    double cb, double cc, int nterms, int nx, int ny, int nz, hipStream_t st);

// projection.hip: divergence (x, y, z) and gradient + cor_vel (z, y, x)
enum { SW_DIV_X = 0, SW_DIV_Y, SW_DIV_Z, SW_GRAD_Z, SW_GRAD_Y, SW_GRAD_X };
bool projection_fusable(const Context &c);
template <typename T>
int launch_projection(const Context &c, int sw, const T *const in[3], T *const out[3], double *const u64[3],
                      float *const u32[3], const float *const du[3], int nx, int ny, int nz, hipStream_t st,
                      double *divp = nullptr, int *divp_blocks = nullptr, const float *const hold[3] = nullptr,
                      int32_t *const res[3] = nullptr);

// ---------------------------------------------------------------- closure network (mlp.hip)
struct MlpDev;
constexpr int MLP_PARAMS = 9414;   // W1 b1 W2 b2 W3 b3 W4 b4 of 10-64-64-64-6
int mlp_alloc(MlpDev *&out);
void mlp_free(MlpDev *M);
int mlp_set_params(MlpDev *M, const float *host, hipStream_t st);
int mlp_get_params(const MlpDev *M, float *host, hipStream_t st);
int mlp_forward(const MlpDev *M, const float *x, float *y, int n, hipStream_t st);
int mlp_train_batch(MlpDev *M, const float *x, const float *t, int n, int apply, float lr, float *loss_host,
                    hipStream_t st);
int mlp_get_grad(const MlpDev *M, float *host, hipStream_t st);
// T_ij of the network (with the Smagorinsky fallback) from the nine gradients
template <typename T>
int launch_closure(const MlpDev *M, const T *const g[9], T *const t[6], size_t n, double gscale, double dfeat,
                   double delta, double cs, double beta, int project, int *nfallback, hipStream_t st);

// ---------------------------------------------------------------- LES fields (les.hip)
// strain S_ij (xx yy zz xy xz yz) and Smagorinsky nut from the nine gradients
template <typename T>
int launch_strain_nut(const T *const g[9], T *const s[6], T *nut, double cs_d2, size_t n, hipStream_t st);
// acc_c += a * x * y (elementwise, FP64 arithmetic)
template <typename T>
int launch_fma3(T *acc, const T *x, const T *y, double a, size_t n, hipStream_t st);
// y += x
template <typename T>
int launch_add(T *y, const T *x, size_t n, hipStream_t st);
// t_ij = 2 nut S_ij
template <typename T>
int launch_smag_stress(const T *nut, const T *const s[6], T *const t[6], size_t n, hipStream_t st);
// top-hat filter of a fine (nx,ny,nz) velocity onto the (nx/f,ny/f,nz/f)
// grid: filtered u_i and tau_ij = bar(u_i u_j) - bar(u_i) bar(u_j)
int launch_box_filter(const double *const u[3], int nx, int ny, int nz, int f, double *const ub[3],
                      double *const tau[6], hipStream_t st);

}  // namespace x3d
