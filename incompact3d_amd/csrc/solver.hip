// Device-resident time step of the incompressible solver (periodic and
// free-slip boxes, and no-slip walls in y as in the channel):
//   boundary_conditions    Case-Channel.f90:153-261 (constant flow rate)
//   calculate_transeq_rhs  transeq.f90:20-591   (skew-symmetric convection + nu lap,
//                                                stretched-y metrics, forcing)
//   int_time (AB3 / RK3)   time_integrators.f90:85-100, 151-157
//   pre_correc             navier.f90:502-789   (free-slip faces, y walls)
//   solve_poisson          navier.f90:27-141    (divergence, poisson, gradp)
//   cor_vel                navier.f90:206-249
// as in the time loop of xcompact3d.f90:29-102.  On grids the fused sweeps
// cover, each direction of the momentum right-hand side is one launch per
// velocity component (AB3 fused into the x sweep) and each direction of the
// divergence and of the gradient + cor_vel is one launch; elsewhere every
// derivative and staggered operator is one line-solve launch whose epilogue
// does the reference's array arithmetic.  precision 0 computes in FP64;
// precision 1 keeps an FP32 copy of the velocity for every derivative, FFT
// and product, while the velocity itself is accumulated in FP64 (AB3 and
// cor_vel) and the Poisson divide and reductions run in FP64.  RK3
// (itimescheme=5), free-slip faces, y walls and stretched y meshes run on
// the per-operator path.
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "x3d_internal.h"

namespace x3d {

template <typename H>
int launch_ab(double *u, const H *f1, const H *f2, const H *f3, double a, double b, double c, int nterms,
              float *shadow, size_t n, hipStream_t st);
template <typename P>
int launch_corvel(double *u, const P *p, float *shadow, size_t n, hipStream_t st);
int launch_cast(const double *a, float *b, size_t n, hipStream_t st);
int launch_widen(const float *a, double *b, size_t n, hipStream_t st);
int launch_split(const double *a, float *hi, int32_t *r, size_t n, hipStream_t st);
int launch_join(const float *hi, const int32_t *r, double *a, size_t n, hipStream_t st);
int launch_tgv(double *ux, double *uy, double *uz, int nx, int ny, int nz, double dx, double dy, double dz,
               hipStream_t st);
template <typename T>
int launch_tgv_sums(const T *ux, const T *uy, const T *uz, const T *const g[9], const T *lx, const T *ly,
                    const T *lz, double xnu, int nx, int ny, int nz, int nxc, int nyc, int nzc, double *partial,
                    double *out4, hipStream_t st);
int launch_zero_face(double *u, float *u32, int nx, int ny, int nz, int dir, int idx, hipStream_t st);
template <typename T>
int launch_max_mean(const T *a, size_t n, double *partial, double *out2, hipStream_t st);
template <typename T>
int launch_absmax_partials(const T *a, size_t n, double *partial, int *nb, hipStream_t st);
int launch_finish_absmax(const double *partial, int nb, size_t n, double *out2, hipStream_t st);
// channel.hip
int launch_channel_init(double *ux, double *uy, double *uz, const double *yc, int nx, int ny, int nz, double noise,
                        unsigned long long seed, hipStream_t st);
int launch_bulk(const double *u, const double *ppy, int nx, int ny, int nz, double coeff, double constant,
                double *partial, double *out2, hipStream_t st);
int launch_cfr_shift(double *u, float *u32, const double *out2, size_t n, hipStream_t st);
template <typename T>
int launch_add_const(T *f, double c, size_t n, hipStream_t st);
template <typename T>
int launch_axpy(T *f, const T *v, double a, size_t n, hipStream_t st);
int launch_wall_velocity(double *const u[3], float *const v[3], const double *dpdx, const double *dpdz, double gdt,
                         int nx, int ny, int nz, int j, hipStream_t st);
template <typename T>
int launch_wall_store(const T *g, double *dst, double gdt, int nx, int ny, int nz, int j, hipStream_t st);

}  // namespace x3d

struct x3d_solver {
  int precision = 0;
  int itime = 0;
  bool fused = false;          // the fused sweeps (momentum.hip, projection.hip) cover this grid
  size_t n = 0;
  int nx = 0, ny = 0, nz = 0;
  double *u64[3] = {nullptr, nullptr, nullptr};   // current velocity (FP64)
  float *u32[3] = {nullptr, nullptr, nullptr};    // its FP32 copy (precision 1)
  double *u64b[3] = {nullptr, nullptr, nullptr};  // next-step buffers (AB3 writes here,
  float *u32b[3] = {nullptr, nullptr, nullptr};   // then the two sets swap)
  // fused FP32-bulk step: the FP64 velocity lives between steps as u32 (its
  // FP32 rounding) plus res (an integer residual), exact (state64.h); u64 is
  // rebuilt from the pair when a caller asks for it (u64_stale)
  bool split = false;
  bool u64_stale = false;
  int32_t *res[3] = {nullptr, nullptr, nullptr};
  void *hist[3][3] = {};       // [component][slot], compute precision
  int slot[3] = {0, 1, 2};     // slot index of f1, f2, f3
  void *w[5] = {};             // work fields, compute precision; after a step
                               // w[0] holds its (last sub-step's) pressure
  bool have_pressure = false;  // w[0] holds a pressure (a step has run)
  // statistics of div u* of the last (sub-)step (navier.f90:348-361): per-block
  // max and sum of |.| pairs, divp_n of them, reduced when asked for
  double *divp = nullptr;
  int divp_n = 0;
  void *diag = nullptr;        // 12 fields for the TGV statistics
  double *w64[4] = {};         // FP64 work for the divergence check of precision 1
  double *partial = nullptr;   // reduction partials
  double *red = nullptr;       // reduction results (device)
  // LES (les_models.f90): 0 none, 1 Smagorinsky (jles=1), 2 learned closure,
  // 3 its eddy-viscosity projection
  int les = 0, iconserv = 0;
  double cs = 0.15, beta = 0.0, gscale = 1.0, delta0 = 1.0;
  x3d::MlpDev *mlp = nullptr;  // not owned
  void *lesw = nullptr;        // 20 fields: g[9], s or T [6], nut, ta, sgs[3]
  int *nfb = nullptr;          // fallback counter of the learned closure (device)
  // no-slip walls in y (ncly1/nclyn = 2): the pressure gradient on each wall
  // over gdt that gradp stores for the next pre_correc, dpdxy1, dpdzy1,
  // dpdxyn, dpdzyn of navier.f90:456-479, nz*nx each, [k*nx + i]
  bool wall1 = false, walln = false;
  double *wallg = nullptr;
  // streamwise forcing of the channel: 0 none, 1 constant flow rate
  // (channel_cfr towards fvalue, Case-Channel.f90:165-171), 2 constant
  // pressure gradient fvalue added to the x right-hand side (cpg, :403-410)
  int forcing = 0;
  double fvalue = 0.0;
  // spin-up rotation of the channel (Case-Channel.f90:413-418): while
  // itime < spinup_time, dux -= wrotation uy and duy += wrotation ux
  double wrotation = 0.0;
  int spinup_time = 0;
  double *ycoord = nullptr;    // wall-normal coordinate of init_channel
};

namespace x3d {
namespace {

// Device copies of the stretched-y metrics (nullptr on a uniform mesh).
enum { M_PPY = 0, M_PPYI, M_PP2Y, M_PP4YY };
template <typename T>
const T *metric(const Context &c, int which) {
  if (c.p.istret == 0) return nullptr;
  if constexpr (sizeof(T) == 8) {
    const double *m[4] = {c.ppy64, c.ppyi64, c.pp2y64, c.pp4yy64};
    return (const T *)m[which];
  } else {
    const float *m[4] = {c.ppy32, c.ppyi32, c.pp2y32, c.pp4yy32};
    return (const T *)m[which];
  }
}

// One line operator with the epilogue t = ac*acc + rc*(mul*result).  On a
// stretched y mesh the y operators multiply by the metric the reference
// passes them: dery and derypv by ppy, deryvp by ppyi (derive.f90:409-417,
// 4572-4578, 4905-4911); deryy and the interpolations by none.  pp, when
// given, replaces that metric.
template <typename T>
int line(const Context &c, int dir, int op, int npaire, const T *u, const T *u2, T *t, const T *acc, const T *mul,
         double ac, double rc, int nx, int ny, int nz, const T *pp = nullptr) {
  LineCall<T> L;
  X3D_TRY(make_line_call<T>(c, dir, op, npaire, &L));
  if (pp) L.pp = pp;
  else if (dir == 1 && (op == OP_DER1 || op == OP_DERPV)) L.pp = metric<T>(c, M_PPY);
  else if (dir == 1 && op == OP_DERVP) L.pp = metric<T>(c, M_PPYI);
  L.nx = nx;
  L.ny = ny;
  L.nz = nz;
  L.u = u;
  L.u2 = u2;
  L.t = t;
  L.acc = acc;
  L.mul = mul;
  L.ac = (T)ac;
  L.rc = (T)rc;
  return launch_line<T>(L, c.stream);
}

// momentum_rhs_eq (transeq.f90:73-470) for one component c of the velocity,
// one line-solve launch per derivative (the path for grids the fused sweeps
// of momentum.hip do not cover):
//   out = xnu*(d2x+d2y+d2z) u_c - 1/2 sum_d [ d_d(u_d u_c) + u_d d_d u_c ]
// accumulated in the reference's order z, y, x.  Parities follow the
// reference calls: the product derivative along d uses the even (p) sets
// when d == c, the velocity derivative when d != c.
template <typename T>
int momentum_component(const Context &c, const T *const u[3], int comp, T *W, T *out, int nx, int ny, int nz) {
  const double xnu = c.xnu;
  const T *uc = u[comp];
  for (int d = 2; d >= 0; --d) {
    const int pp = d == comp ? 1 : 0;   // parity of the product u_d*u_c
    const int pv = d == comp ? 0 : 1;   // parity of u_c along d
    if (d == 2) {
      X3D_TRY(line<T>(c, d, OP_DER1, pp, u[d], uc, W, nullptr, nullptr, 0.0, 1.0, nx, ny, nz));
      X3D_TRY(line<T>(c, d, OP_DER1, pv, uc, nullptr, W, W, u[d], 1.0, 1.0, nx, ny, nz));
      X3D_TRY(line<T>(c, d, OP_DER2, pv, uc, nullptr, W, W, nullptr, -0.5, xnu, nx, ny, nz));
    } else {
      X3D_TRY(line<T>(c, d, OP_DER1, pp, u[d], uc, W, W, nullptr, 1.0, -0.5, nx, ny, nz));
      X3D_TRY(line<T>(c, d, OP_DER1, pv, uc, nullptr, W, W, u[d], 1.0, -0.5, nx, ny, nz));
      T *dst = d == 0 ? out : W;
      if (d == 1 && c.p.istret != 0) {
        // transeq.f90:337-370: nu (pp2y deryy u - pp4y dery u), dery with ppy
        X3D_TRY(line<T>(c, d, OP_DER2, pv, uc, nullptr, W, W, nullptr, 1.0, xnu, nx, ny, nz, metric<T>(c, M_PP2Y)));
        X3D_TRY(line<T>(c, d, OP_DER1, pv, uc, nullptr, W, W, nullptr, 1.0, -xnu, nx, ny, nz, metric<T>(c, M_PP4YY)));
      } else {
        X3D_TRY(line<T>(c, d, OP_DER2, pv, uc, nullptr, dst, W, nullptr, 1.0, xnu, nx, ny, nz));
      }
    }
  }
  return X3D_OK;
}

// ---------------------------------------------------------------- LES (les_models.f90)
// SGS force of the current velocity into sgs[3] (compute precision):
// Smagorinsky in the nonconservative form (sgs_mom_nonconservative,
// les_models.f90:1091-1250) or the conservative one (:1327-1456), or the
// learned closure T_ij followed by d_j T_ij with the conservative parities.
template <typename T>
int stress_divergence(const Context &c, T *const t[6], T *const sgs[3], int nx, int ny, int nz) {
  // les_models.f90:1384-1440; t = xx yy zz xy xz yz
  X3D_TRY(line<T>(c, 0, OP_DER1, 0, t[0], nullptr, sgs[0], nullptr, nullptr, 0, 1, nx, ny, nz));
  X3D_TRY(line<T>(c, 0, OP_DER1, 0, t[3], nullptr, sgs[1], nullptr, nullptr, 0, 1, nx, ny, nz));
  X3D_TRY(line<T>(c, 0, OP_DER1, 0, t[4], nullptr, sgs[2], nullptr, nullptr, 0, 1, nx, ny, nz));
  X3D_TRY(line<T>(c, 1, OP_DER1, 0, t[3], nullptr, sgs[0], sgs[0], nullptr, 1, 1, nx, ny, nz));
  X3D_TRY(line<T>(c, 1, OP_DER1, 1, t[1], nullptr, sgs[1], sgs[1], nullptr, 1, 1, nx, ny, nz));
  X3D_TRY(line<T>(c, 1, OP_DER1, 0, t[5], nullptr, sgs[2], sgs[2], nullptr, 1, 1, nx, ny, nz));
  X3D_TRY(line<T>(c, 2, OP_DER1, 0, t[4], nullptr, sgs[0], sgs[0], nullptr, 1, 1, nx, ny, nz));
  X3D_TRY(line<T>(c, 2, OP_DER1, 0, t[5], nullptr, sgs[1], sgs[1], nullptr, 1, 1, nx, ny, nz));
  X3D_TRY(line<T>(c, 2, OP_DER1, 0, t[2], nullptr, sgs[2], sgs[2], nullptr, 1, 1, nx, ny, nz));
  return X3D_OK;
}

template <typename T>
int compute_sgs(const Context &c, x3d_solver *S, const T *const u[3], T *const sgs[3]) {
  const int nx = S->nx, ny = S->ny, nz = S->nz;
  const size_t n = S->n;
  T *f = (T *)S->lesw;
  T *g[9], *s6[6];
  for (int i = 0; i < 9; ++i) g[i] = f + i * n;
  for (int i = 0; i < 6; ++i) s6[i] = f + (9 + i) * n;
  T *nut = f + 15 * n, *ta = f + 16 * n;
  // the nine gradients with the parities of smag (les_models.f90:210-238):
  // g[3*d + comp] = d u_comp / d x_d, parity 0 when d == comp
  for (int d = 0; d < 3; ++d)
    for (int comp = 0; comp < 3; ++comp)
      X3D_TRY(line<T>(c, d, OP_DER1, d == comp ? 0 : 1, u[comp], nullptr, g[3 * d + comp], nullptr, nullptr, 0, 1,
                      nx, ny, nz));
  const T *cg[9];
  for (int i = 0; i < 9; ++i) cg[i] = g[i];
  const double delta = std::cbrt(c.dir[0].d * c.dir[1].d * c.dir[2].d);   // les_models.f90:36-39
  if (S->les >= 2) {
    X3D_TRY(launch_closure<T>(S->mlp, cg, s6, n, S->gscale, delta / S->delta0, delta, S->cs, S->beta,
                              S->les == 3 ? 1 : 0, S->nfb, c.stream));
    return stress_divergence<T>(c, s6, sgs, nx, ny, nz);
  }
  X3D_TRY(launch_strain_nut<T>(cg, s6, nut, (S->cs * delta) * (S->cs * delta), n, c.stream));
  if (S->iconserv == 1) {
    X3D_TRY(launch_smag_stress<T>(nut, s6, s6, n, c.stream));
    return stress_divergence<T>(c, s6, sgs, nx, ny, nz);
  }
  // s6 = xx yy zz xy xz yz; S_{comp,d} for the 2 S_ij d_j nut term
  const int sidx[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
  for (int d = 0; d < 3; ++d) {
    X3D_TRY(line<T>(c, d, OP_DER1, 1, nut, nullptr, ta, nullptr, nullptr, 0, 1, nx, ny, nz));
    for (int comp = 0; comp < 3; ++comp) {
      // sgs = sgs + nut * d2_d u (d == x: sgs = d2 u * nut)
      X3D_TRY(line<T>(c, d, OP_DER2, d == comp ? 0 : 1, u[comp], nullptr, sgs[comp], d == 0 ? nullptr : sgs[comp],
                      nut, d == 0 ? 0.0 : 1.0, 1.0, nx, ny, nz));
      X3D_TRY(launch_fma3<T>(sgs[comp], s6[sidx[comp][d]], ta, 2.0, n, c.stream));
    }
  }
  return X3D_OK;
}

// divergence (navier.f90:257-372), result (pressure grid) in W[0]
template <typename T>
int divergence(const Context &c, const T *const u[3], T *const W[4], int nx, int ny, int nz) {
  const int nxm = c.nxm, nym = c.nym;
  X3D_TRY(line<T>(c, 0, OP_DERVP, 0, u[0], nullptr, W[0], nullptr, nullptr, 0, 1, nx, ny, nz));
  X3D_TRY(line<T>(c, 0, OP_INTERVP, 1, u[1], nullptr, W[1], nullptr, nullptr, 0, 1, nx, ny, nz));
  X3D_TRY(line<T>(c, 0, OP_INTERVP, 1, u[2], nullptr, W[2], nullptr, nullptr, 0, 1, nx, ny, nz));
  X3D_TRY(line<T>(c, 1, OP_INTERVP, 1, W[0], nullptr, W[3], nullptr, nullptr, 0, 1, nxm, ny, nz));
  X3D_TRY(line<T>(c, 1, OP_DERVP, 0, W[1], nullptr, W[3], W[3], nullptr, 1, 1, nxm, ny, nz));
  X3D_TRY(line<T>(c, 1, OP_INTERVP, 1, W[2], nullptr, W[1], nullptr, nullptr, 0, 1, nxm, ny, nz));
  X3D_TRY(line<T>(c, 2, OP_INTERVP, 1, W[3], nullptr, W[0], nullptr, nullptr, 0, 1, nxm, nym, nz));
  X3D_TRY(line<T>(c, 2, OP_DERVP, 0, W[1], nullptr, W[0], W[0], nullptr, 1, 1, nxm, nym, nz));
  return X3D_OK;
}

// gradp (navier.f90:386-431) with cor_vel fused: u_i <- u_i - grad_i p.
// pp3 (pressure grid) in W[0]; for FP32 compute, or with walls in y, the
// gradient goes to W[2] and is subtracted from the FP64 velocity; on the
// walls its x and z components over gdt are kept for pre_correc
// (navier.f90:456-479).
template <typename T>
int gradp_correct(const Context &c, x3d_solver *S, T *const u[3], T *const W[5], double gdt, bool apply = true) {
  const int nx = S->nx, ny = S->ny, nz = S->nz, nxm = c.nxm, nym = c.nym, nzm = c.nzm;
  X3D_TRY(line<T>(c, 2, OP_INTERPV, 1, W[0], nullptr, W[1], nullptr, nullptr, 0, 1, nxm, nym, nzm));
  X3D_TRY(line<T>(c, 2, OP_DERPV, 1, W[0], nullptr, W[2], nullptr, nullptr, 0, 1, nxm, nym, nzm));
  X3D_TRY(line<T>(c, 1, OP_INTERPV, 1, W[1], nullptr, W[3], nullptr, nullptr, 0, 1, nxm, nym, nz));
  X3D_TRY(line<T>(c, 1, OP_DERPV, 1, W[1], nullptr, W[4], nullptr, nullptr, 0, 1, nxm, nym, nz));
  X3D_TRY(line<T>(c, 1, OP_INTERPV, 1, W[2], nullptr, W[1], nullptr, nullptr, 0, 1, nxm, nym, nz));
  // px from W[3], py from W[4], pz from W[1]; the pressure W[0] is kept
  const T *src[3] = {W[3], W[4], W[1]};
  const int ops[3] = {OP_DERPV, OP_INTERPV, OP_INTERPV};
  const bool walls = S->wall1 || S->walln;
  const size_t np = (size_t)nx * nz;
  for (int i = 0; i < 3; ++i) {
    if (!apply && !(walls && i != 1)) continue;
    if (sizeof(T) == 8 && !walls) {
      X3D_TRY(line<T>(c, 0, ops[i], 1, src[i], nullptr, u[i], u[i], nullptr, 1.0, -1.0, nxm, ny, nz));
      continue;
    }
    X3D_TRY(line<T>(c, 0, ops[i], 1, src[i], nullptr, W[2], nullptr, nullptr, 0, 1, nxm, ny, nz));
    if (walls && i != 1) {
      // dpdxy1 / dpdzy1 (j = 1) and dpdxyn / dpdzyn (j = ny)
      const int q = i == 0 ? 0 : 1;
      if (S->wall1) X3D_TRY(launch_wall_store<T>(W[2], S->wallg + q * np, gdt, nx, ny, nz, 0, c.stream));
      if (S->walln) X3D_TRY(launch_wall_store<T>(W[2], S->wallg + (2 + q) * np, gdt, nx, ny, nz, ny - 1, c.stream));
    }
    if (!apply) continue;
    X3D_TRY(launch_corvel<T>(S->u64[i], W[2], sizeof(T) == 8 ? nullptr : S->u32[i], S->n, c.stream));
  }
  return X3D_OK;
}

// The six y/z momentum sweeps of step_once (no SGS force), nrep times, into
// the work buffer W[0]; *bytes = algorithmic bytes per launch on average.
template <typename T>
int momentum_yz(Context &c, x3d_solver *S, int nrep, double *bytes) {
  const int nx = S->nx, ny = S->ny, nz = S->nz;
  const T *cu[3];
  for (int i = 0; i < 3; ++i) cu[i] = sizeof(T) == 8 ? (const T *)S->u64[i] : (const T *)S->u32[i];
  T *W0 = (T *)S->w[0];
  double fields = 0;
  for (int comp = 0; comp < 3; ++comp) {
    fields += (comp == 2 ? 1 : 2) + 1;        // z: u_z (and u_c), out
    fields += (comp == 1 ? 1 : 2) + 2;        // y: u_y (and u_c), acc, out
  }
  *bytes = fields / 6.0 * (double)S->n * sizeof(T);
  S->have_pressure = false;   // the probe writes W[0]
  for (int r = 0; r < nrep; ++r) {
    for (int comp = 0; comp < 3; ++comp) {
      X3D_TRY(launch_momentum<T>(c, 2, comp, cu, nullptr, W0, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, 0,
                                 0, nx, ny, nz, c.stream));
      X3D_TRY(launch_momentum<T>(c, 1, comp, cu, W0, W0, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, 0, 0,
                                 nx, ny, nz, c.stream));
    }
  }
  return X3D_OK;
}

// pre_correc (navier.f90:502-789) for the boundaries this solver covers:
// zero normal velocity on free-slip faces (nothing for periodic ones); on a
// no-slip y wall (navier.f90:616-690) u = dpdxy gdt, v = 0, w = dpdzy gdt,
// the wall velocity that the pressure correction of this sub-step brings
// back to zero.
int pre_correc(const Context &c, x3d_solver *S, double gdt) {
  float *const u32[3] = {S->u32[0], S->u32[1], S->u32[2]};
  const int nn[3] = {S->nx, S->ny, S->nz};
  for (int d = 0; d < 3; ++d) {
    if (c.dir[d].bc1 == 1) X3D_TRY(launch_zero_face(S->u64[d], u32[d], S->nx, S->ny, S->nz, d, 0, c.stream));
    if (c.dir[d].bcn == 1)
      X3D_TRY(launch_zero_face(S->u64[d], u32[d], S->nx, S->ny, S->nz, d, nn[d] - 1, c.stream));
  }
  const size_t np = (size_t)S->nx * S->nz;
  double *const u64[3] = {S->u64[0], S->u64[1], S->u64[2]};
  if (S->wall1)
    X3D_TRY(launch_wall_velocity(u64, u32, S->wallg, S->wallg + np, gdt, S->nx, S->ny, S->nz, 0, c.stream));
  if (S->walln)
    X3D_TRY(launch_wall_velocity(u64, u32, S->wallg + 2 * np, S->wallg + 3 * np, gdt, S->nx, S->ny, S->nz, S->ny - 1,
                                 c.stream));
  return X3D_OK;
}

// boundary_conditions of the channel (Case-Channel.f90:165-171): the
// constant flow rate of channel_cfr, u <- u + (constant - ub) with
// ub = dy / (yly nx nz) sum u / ppy (Case-Channel.f90:220-261)
int channel_cfr(const Context &c, x3d_solver *S) {
  if (S->forcing != 1) return X3D_OK;
  const double coeff = c.dir[1].d / (c.p.yly * (double)((long long)S->nx * S->nz));
  X3D_TRY(launch_bulk(S->u64[0], c.p.istret != 0 ? c.ppy64 : nullptr, S->nx, S->ny, S->nz, coeff, S->fvalue,
                      S->partial, S->red, c.stream));
  return launch_cfr_shift(S->u64[0], S->u32[0], S->red, S->n, c.stream);
}

// momentum_forcing_channel (Case-Channel.f90:396-420) on component comp's
// right-hand side f1: the constant pressure gradient dux1 += fcpg (cpg,
// :403-410), then the spin-up rotation while itime < spinup_time (:413-418),
// dux1 -= wrotation uy1 and duy1 += wrotation ux1 with the sub-step's velocity
template <typename T>
int channel_forcing(const Context &c, x3d_solver *S, int comp, T *f1, const T *const u[3]) {
  if (comp == 0 && S->forcing == 2) X3D_TRY(launch_add_const<T>(f1, S->fvalue, S->n, c.stream));
  if (S->itime < S->spinup_time && comp < 2)
    X3D_TRY(launch_axpy<T>(f1, u[1 - comp], comp == 0 ? -S->wrotation : S->wrotation, S->n, c.stream));
  return X3D_OK;
}

// One RK3 step (itimescheme=5, xcompact3d.f90:46-87): three sub-steps of
// transeq RHS, the low-storage update of time_integrators.f90:151-157 with
// the coefficients of variables.f90:1388-1399 (FP64 accumulation),
// pre_correc, divergence, Poisson, gradp + cor_vel.  hist slot[0] holds f1,
// slot[1] the previous sub-step's f1.
template <typename T>
int step_rk3(Context &c, x3d_solver *S) {
  const int nx = S->nx, ny = S->ny, nz = S->nz;
  T *u[3];
  for (int i = 0; i < 3; ++i) u[i] = sizeof(T) == 8 ? (T *)S->u64[i] : (T *)S->u32[i];
  T *W[5] = {(T *)S->w[0], (T *)S->w[1], (T *)S->w[2], (T *)S->w[3], (T *)S->w[4]};
  S->itime += 1;
  const double dt = c.p.dt;
  const double adt[3] = {(8.0 / 15.0) * dt, (5.0 / 12.0) * dt, (3.0 / 4.0) * dt};
  const double bdt[3] = {0.0, (-17.0 / 60.0) * dt, (-5.0 / 12.0) * dt};
  const double gdt[3] = {adt[0], adt[1] + bdt[1], adt[2] + bdt[2]};
  const T *cu[3] = {u[0], u[1], u[2]};
  for (int itr = 0; itr < 3; ++itr) {
    X3D_TRY(channel_cfr(c, S));
    T *sgs[3] = {nullptr, nullptr, nullptr};
    if (S->les) {
      for (int i = 0; i < 3; ++i) sgs[i] = (T *)S->lesw + (17 + i) * S->n;
      X3D_TRY(compute_sgs<T>(c, S, cu, sgs));
    }
    for (int comp = 0; comp < 3; ++comp) {
      T *f1 = (T *)S->hist[comp][S->slot[0]];
      X3D_TRY(momentum_component<T>(c, cu, comp, W[0], f1, nx, ny, nz));
      if (S->les) X3D_TRY(launch_add<T>(f1, sgs[comp], S->n, c.stream));
      X3D_TRY(channel_forcing<T>(c, S, comp, f1, cu));
    }
    for (int comp = 0; comp < 3; ++comp) {
      const T *f1 = (T *)S->hist[comp][S->slot[0]];
      const T *f2 = (T *)S->hist[comp][S->slot[1]];
      // itr 1: u += gdt(1) f1 with gdt(1) = adt(1); then u += adt f1 + bdt f2
      X3D_TRY(launch_ab<T>(S->u64[comp], f1, f2, nullptr, adt[itr], bdt[itr], 0.0, itr == 0 ? 1 : 2,
                           sizeof(T) == 8 ? nullptr : S->u32[comp], S->n, c.stream));
    }
    std::swap(S->slot[0], S->slot[1]);   // f2 <- f1
    X3D_TRY(pre_correc(c, S, gdt[itr]));
    X3D_TRY(divergence<T>(c, cu, W, nx, ny, nz));
    X3D_TRY(launch_absmax_partials<T>(W[0], (size_t)c.nxm * c.nym * c.nzm, S->divp, &S->divp_n, c.stream));
    X3D_TRY(poisson_solve<T>(c, W[0], c.stream));
    X3D_TRY(gradp_correct<T>(c, S, u, W, gdt[itr]));
  }
  return X3D_OK;
}

template <typename T>
int step_once(Context &c, x3d_solver *S) {
  const int nx = S->nx, ny = S->ny, nz = S->nz;
  T *u[3];
  for (int i = 0; i < 3; ++i) u[i] = sizeof(T) == 8 ? (T *)S->u64[i] : (T *)S->u32[i];
  T *W[5] = {(T *)S->w[0], (T *)S->w[1], (T *)S->w[2], (T *)S->w[3], (T *)S->w[4]};
  S->itime += 1;
  // AB3 coefficients (time_integrators.f90:85-100; variables.f90:1365-1375)
  const double dt = c.p.dt;
  double a, b = 0, cc = 0;
  int nterms;
  if (S->itime == 1) {
    a = dt; nterms = 1;
  } else if (S->itime == 2) {
    a = 1.5 * dt; b = -(0.5 * dt); nterms = 2;
  } else {
    a = (23.0 / 12.0) * dt; b = -(16.0 / 12.0) * dt; cc = (5.0 / 12.0) * dt; nterms = 3;
  }
  // gdt(1) of variables.f90:1368-1372, the same at every step
  const double gdt = (23.0 / 12.0) * dt + (-(16.0 / 12.0) * dt) + (5.0 / 12.0) * dt;
  X3D_TRY(channel_cfr(c, S));
  // transeq (z, y, x sweeps per component) with AB3 fused into the x sweep;
  // the new velocity goes to the second buffer set
  const T *cu[3] = {u[0], u[1], u[2]};
  // SGS force (transeq.f90:498-503), added to the RHS before AB3
  T *sgs[3] = {nullptr, nullptr, nullptr};
  if (S->les) {
    for (int i = 0; i < 3; ++i) sgs[i] = (T *)S->lesw + (17 + i) * S->n;
    X3D_TRY(compute_sgs<T>(c, S, cu, sgs));
  }
  if (S->fused) {
    for (int comp = 0; comp < 3; ++comp) {
      T *f1 = (T *)S->hist[comp][S->slot[0]];
      const T *f2 = (T *)S->hist[comp][S->slot[1]];
      const T *f3 = (T *)S->hist[comp][S->slot[2]];
      X3D_TRY(launch_momentum<T>(c, 2, comp, cu, sgs[comp], W[0], nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0,
                                 0, 0, nx, ny, nz, c.stream));
      X3D_TRY(launch_momentum<T>(c, 1, comp, cu, W[0], W[0], nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, 0,
                                 0, nx, ny, nz, c.stream));
      // FP64: the new velocity into u64b; FP32: the increment into u64b (as
      // FP32) and the new FP32 velocity into u32b, u64 updated by GRAD_X
      if constexpr (sizeof(T) == 8)
        X3D_TRY(launch_momentum<T>(c, 0, comp, cu, W[0], f1, f2, f3, S->u64b[comp], nullptr, nullptr, a, b, cc,
                                   nterms, nx, ny, nz, c.stream));
      else
        X3D_TRY(launch_momentum<T>(c, 0, comp, cu, W[0], f1, f2, f3, nullptr, (float *)S->u64b[comp],
                                   S->u32b[comp], a, b, cc, nterms, nx, ny, nz, c.stream));
    }
    for (int i = 0; i < 3; ++i) {
      if (sizeof(T) == 8) std::swap(S->u64[i], S->u64b[i]);
      else std::swap(S->u32[i], S->u32b[i]);
      u[i] = sizeof(T) == 8 ? (T *)S->u64[i] : (T *)S->u32[i];
      cu[i] = u[i];
    }
  } else {
    for (int comp = 0; comp < 3; ++comp) {
      T *f1 = (T *)S->hist[comp][S->slot[0]];
      X3D_TRY(momentum_component<T>(c, cu, comp, W[0], f1, nx, ny, nz));
      if (S->les) X3D_TRY(launch_add<T>(f1, sgs[comp], S->n, c.stream));
      X3D_TRY(channel_forcing<T>(c, S, comp, f1, cu));
    }
    for (int comp = 0; comp < 3; ++comp) {
      const T *f1 = (T *)S->hist[comp][S->slot[0]];
      const T *f2 = (T *)S->hist[comp][S->slot[1]];
      const T *f3 = (T *)S->hist[comp][S->slot[2]];
      X3D_TRY(launch_ab<T>(S->u64[comp], f1, f2, f3, a, b, cc, nterms, sizeof(T) == 8 ? nullptr : S->u32[comp],
                           S->n, c.stream));
    }
  }
  // history rotation (f3 <- f2 <- f1) by slot permutation
  const int s1 = S->slot[0], s2 = S->slot[1], s3 = S->slot[2];
  S->slot[0] = s3;
  S->slot[1] = s1;
  S->slot[2] = s2;
  // projection
  if (S->fused) {
    T *const onone[3] = {nullptr, nullptr, nullptr};
    const T *const dx_in[3] = {cu[0], cu[1], cu[2]};
    T *const dx_out[3] = {W[0], W[1], W[2]};
    X3D_TRY(launch_projection<T>(c, SW_DIV_X, dx_in, dx_out, nullptr, nullptr, nullptr, nx, ny, nz, c.stream));
    const T *const dy_in[3] = {W[0], W[1], W[2]};
    T *const dy_out[3] = {W[3], W[1], nullptr};
    X3D_TRY(launch_projection<T>(c, SW_DIV_Y, dy_in, dy_out, nullptr, nullptr, nullptr, nx, ny, nz, c.stream));
    const T *const dz_in[3] = {W[3], W[1], nullptr};
    T *const dz_out[3] = {W[0], nullptr, nullptr};
    X3D_TRY(launch_projection<T>(c, SW_DIV_Z, dz_in, dz_out, nullptr, nullptr, nullptr, nx, ny, nz, c.stream,
                                 S->divp, &S->divp_n));
    X3D_TRY(poisson_solve<T>(c, W[0], c.stream));
    const T *const gz_in[3] = {W[0], nullptr, nullptr};
    T *const gz_out[3] = {W[1], W[2], nullptr};
    X3D_TRY(launch_projection<T>(c, SW_GRAD_Z, gz_in, gz_out, nullptr, nullptr, nullptr, nx, ny, nz, c.stream));
    const T *const gy_in[3] = {W[1], W[2], nullptr};
    T *const gy_out[3] = {W[3], W[4], W[1]};   // W[0] keeps the pressure
    X3D_TRY(launch_projection<T>(c, SW_GRAD_Y, gy_in, gy_out, nullptr, nullptr, nullptr, nx, ny, nz, c.stream));
    const T *const gx_in[3] = {W[3], W[4], W[1]};
    double *const u64[3] = {S->u64[0], S->u64[1], S->u64[2]};
    float *const u32[3] = {S->u32[0], S->u32[1], S->u32[2]};
    const float *const du[3] = {(const float *)S->u64b[0], (const float *)S->u64b[1], (const float *)S->u64b[2]};
    if (S->split) {
      // after the swap u32b holds the rounding of the velocity before the step
      const float *const hold[3] = {S->u32b[0], S->u32b[1], S->u32b[2]};
      X3D_TRY(launch_projection<T>(c, SW_GRAD_X, gx_in, onone, nullptr, u32, du, nx, ny, nz, c.stream, nullptr,
                                   nullptr, hold, S->res));
      S->u64_stale = true;
    } else {
      X3D_TRY(launch_projection<T>(c, SW_GRAD_X, gx_in, onone, u64, sizeof(T) == 8 ? nullptr : u32,
                                   sizeof(T) == 8 ? nullptr : du, nx, ny, nz, c.stream));
    }
  } else {
    X3D_TRY(pre_correc(c, S, gdt));
    X3D_TRY(divergence<T>(c, cu, W, nx, ny, nz));
    X3D_TRY(launch_absmax_partials<T>(W[0], (size_t)c.nxm * c.nym * c.nzm, S->divp, &S->divp_n, c.stream));
    X3D_TRY(poisson_solve<T>(c, W[0], c.stream));
    X3D_TRY(gradp_correct<T>(c, S, u, W, gdt));
  }
  return X3D_OK;
}

template <typename T>
int tgv_stats(Context &c, x3d_solver *S, double out4[4]) {
  const int nx = S->nx, ny = S->ny, nz = S->nz;
  const size_t n = S->n;
  if (!S->diag) X3D_HIP(hipMalloc(&S->diag, 12 * n * sizeof(T)));
  T *g[12];
  for (int i = 0; i < 12; ++i) g[i] = (T *)S->diag + i * n;
  const T *u[3];
  for (int i = 0; i < 3; ++i) u[i] = sizeof(T) == 8 ? (const T *)S->u64[i] : (const T *)S->u32[i];
  // (dux/dx, duy/dx, duz/dx, dux/dy, ..., duz/dz) with the parities of Case-TGV.f90:262-270
  for (int d = 0; d < 3; ++d)
    for (int comp = 0; comp < 3; ++comp) {
      const int pv = d == comp ? 0 : 1;
      X3D_TRY(line<T>(c, d, OP_DER1, pv, u[comp], nullptr, g[3 * d + comp], nullptr, nullptr, 0, 1, nx, ny, nz));
    }
  // laplacian per component: d2x + d2y + d2z (Case-TGV.f90:330-360)
  for (int comp = 0; comp < 3; ++comp)
    for (int d = 0; d < 3; ++d) {
      const int pv = d == comp ? 0 : 1;
      T *lap = g[9 + comp];
      X3D_TRY(line<T>(c, d, OP_DER2, pv, u[comp], nullptr, lap, d == 0 ? nullptr : lap, nullptr, 1, 1, nx, ny, nz));
    }
  const T *gg[9];
  for (int i = 0; i < 9; ++i) gg[i] = g[i];
  // free-slip directions count nm = n-1 cells (Case-TGV.f90:213-242)
  const int nxc = c.dir[0].bc1 == 1 ? c.nxm : nx, nyc = c.dir[1].bc1 == 1 ? c.nym : ny,
            nzc = c.dir[2].bc1 == 1 ? c.nzm : nz;
  X3D_TRY(launch_tgv_sums<T>(u[0], u[1], u[2], gg, g[9], g[10], g[11], c.xnu, nx, ny, nz, nxc, nyc, nzc, S->partial,
                             S->red, c.stream));
  double h[4];
  X3D_HIP(hipMemcpyAsync(h, S->red, 4 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  X3D_HIP(hipStreamSynchronize(c.stream));
  const double ncell = (double)nxc * nyc * nzc;
  out4[0] = h[0] / ncell;   // eek
  out4[1] = h[1] / ncell;   // eps
  out4[2] = h[3] / ncell;   // eps2
  out4[3] = h[2] / ncell;   // enst
  return X3D_OK;
}

int alloc_solver(Context &c, x3d_solver *S) {
  const size_t n = S->n;
  const size_t es = S->precision == 0 ? 8 : 4;
  for (int i = 0; i < 3; ++i) {
    X3D_HIP(hipMalloc(&S->u64[i], n * 8));
    if (S->precision == 1) X3D_HIP(hipMalloc(&S->u32[i], n * 4));
    if (S->fused) {
      X3D_HIP(hipMalloc(&S->u64b[i], n * 8));
      if (S->precision == 1) X3D_HIP(hipMalloc(&S->u32b[i], n * 4));
    }
    if (S->split) X3D_HIP(hipMalloc(&S->res[i], n * 4));
    for (int j = 0; j < 3; ++j) {
      X3D_HIP(hipMalloc(&S->hist[i][j], n * es));
      X3D_HIP(hipMemsetAsync(S->hist[i][j], 0, n * es, c.stream));
    }
  }
  const size_t np = (size_t)c.nxm * c.nym * c.nzm;
  const size_t nw = n > np ? n : np;
  for (int i = 0; i < 5; ++i) X3D_HIP(hipMalloc(&S->w[i], nw * es));
  X3D_HIP(hipMalloc(&S->partial, 8 * 1024 * sizeof(double)));
  // div u* partials: RED_BLOCKS pairs (per-operator step) or one pair per
  // block of the fused z divergence sweep (at most nx * ny / 16 blocks)
  X3D_HIP(hipMalloc(&S->divp, 2 * ((size_t)S->nx * S->ny + 1024) * sizeof(double)));
  X3D_HIP(hipMalloc(&S->red, 8 * sizeof(double)));
  if (S->wall1 || S->walln) X3D_HIP(hipMalloc(&S->wallg, 4 * (size_t)S->nx * S->nz * sizeof(double)));
  X3D_HIP(hipStreamSynchronize(c.stream));
  return X3D_OK;
}

void free_solver(x3d_solver *S) {
  for (int i = 0; i < 3; ++i) {
    if (S->u64[i]) (void)hipFree(S->u64[i]);
    if (S->u32[i]) (void)hipFree(S->u32[i]);
    if (S->u64b[i]) (void)hipFree(S->u64b[i]);
    if (S->u32b[i]) (void)hipFree(S->u32b[i]);
    if (S->res[i]) (void)hipFree(S->res[i]);
    for (int j = 0; j < 3; ++j)
      if (S->hist[i][j]) (void)hipFree(S->hist[i][j]);
  }
  for (int i = 0; i < 5; ++i)
    if (S->w[i]) (void)hipFree(S->w[i]);
  for (int i = 0; i < 4; ++i)
    if (S->w64[i]) (void)hipFree(S->w64[i]);
  if (S->diag) (void)hipFree(S->diag);
  if (S->lesw) (void)hipFree(S->lesw);
  if (S->nfb) (void)hipFree(S->nfb);
  if (S->partial) (void)hipFree(S->partial);
  if (S->divp) (void)hipFree(S->divp);
  if (S->red) (void)hipFree(S->red);
  if (S->wallg) (void)hipFree(S->wallg);
  if (S->ycoord) (void)hipFree(S->ycoord);
}

// u64 was just written: refresh its FP32 copy (and, in split form, the
// residual)
int narrow_velocity(Context &c, x3d_solver *S) {
  for (int i = 0; i < 3; ++i) {
    if (S->split) X3D_TRY(launch_split(S->u64[i], S->u32[i], S->res[i], S->n, c.stream));
    else if (S->precision == 1) X3D_TRY(launch_cast(S->u64[i], S->u32[i], S->n, c.stream));
  }
  S->u64_stale = false;
  return X3D_OK;
}

// u64 is about to be read: rebuild it from the split form if a step changed it
int sync_velocity(Context &c, x3d_solver *S) {
  if (!S->split || !S->u64_stale) return X3D_OK;
  for (int i = 0; i < 3; ++i) X3D_TRY(launch_join(S->u32[i], S->res[i], S->u64[i], S->n, c.stream));
  S->u64_stale = false;
  return X3D_OK;
}

int reset_history(Context &c, x3d_solver *S) {
  // case.f90:131-136 seeds the history with u; step 1 (Euler) and step 2
  // (AB2) never read the seeded values, so the slots only need defined data.
  const size_t es = S->precision == 0 ? 8 : 4;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) X3D_HIP(hipMemsetAsync(S->hist[i][j], 0, S->n * es, c.stream));
  S->slot[0] = 0;
  S->slot[1] = 1;
  S->slot[2] = 2;
  S->itime = 0;
  S->have_pressure = false;
  S->divp_n = 0;
  // dpdxy1 .. dpdzyn start at zero (the wall velocity of the first pre_correc)
  if (S->wallg) X3D_HIP(hipMemsetAsync(S->wallg, 0, 4 * (size_t)S->nx * S->nz * sizeof(double), c.stream));
  X3D_TRY(narrow_velocity(c, S));
  X3D_HIP(hipStreamSynchronize(c.stream));
  return X3D_OK;
}

}  // namespace
}  // namespace x3d

using namespace x3d;

extern "C" {

int x3d_solver_create(int precision, x3d_solver **out) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!out) return fail(X3D_ERR_ARG, "null output");
  if (precision != 0 && precision != 1) return fail(X3D_ERR_ARG, "precision must be 0 (FP64) or 1 (FP32 bulk)");
  for (int d = 0; d < 3; ++d) {
    const int b1 = c->dir[d].bc1, bn = c->dir[d].bcn;
    if (d == 1 && b1 != 0 && bn != 0) continue;   // free-slip or no-slip faces in y, in any pairing
    if (!((b1 == 0 && bn == 0) || (b1 == 1 && bn == 1)))
      return fail(X3D_ERR_UNSUPPORTED,
                  "the device time step covers periodic and free-slip (ncl=1) boundaries, and no-slip walls "
                  "(ncl=2) in y only");
  }
  if (!c->poisson) return fail(X3D_ERR_UNSUPPORTED, "no Poisson solver for this grid");
  if (c->p.itimescheme != 3 && c->p.itimescheme != 5)
    return fail(X3D_ERR_UNSUPPORTED, "the device time step implements AB3 (itimescheme=3) and RK3 (itimescheme=5)");
  auto *S = new x3d_solver();
  S->precision = precision;
  S->nx = c->p.nx;
  S->ny = c->p.ny;
  S->nz = c->p.nz;
  S->n = (size_t)S->nx * S->ny * S->nz;
  S->wall1 = c->dir[1].bc1 == 2;
  S->walln = c->dir[1].bcn == 2;
  S->fused = c->p.itimescheme == 3 && c->p.istret == 0 && momentum_fusable(*c) && projection_fusable(*c) &&
             getenv("X3D_UNFUSED_MOMENTUM") == nullptr;
  S->split = S->fused && precision == 1;
  int r = alloc_solver(*c, S);
  if (r != X3D_OK) {
    free_solver(S);
    delete S;
    return r;
  }
  *out = S;
  return X3D_OK;
}

int x3d_solver_destroy(x3d_solver *S) {
  if (!S) return X3D_OK;
  (void)hipDeviceSynchronize();
  free_solver(S);
  delete S;
  return X3D_OK;
}

int x3d_solver_init_tgv(x3d_solver *S) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S) return fail(X3D_ERR_ARG, "null solver");
  X3D_TRY(launch_tgv(S->u64[0], S->u64[1], S->u64[2], S->nx, S->ny, S->nz, c->dir[0].d, c->dir[1].d, c->dir[2].d,
                     c->stream));
  return reset_history(*c, S);
}

int x3d_solver_set_velocity(x3d_solver *S, const double *ux, const double *uy, const double *uz) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || !ux || !uy || !uz) return fail(X3D_ERR_ARG, "null pointer");
  const double *src[3] = {ux, uy, uz};
  for (int i = 0; i < 3; ++i)
    X3D_HIP(hipMemcpyAsync(S->u64[i], src[i], S->n * 8, hipMemcpyDeviceToDevice, c->stream));
  return reset_history(*c, S);
}

int x3d_solver_get_velocity(x3d_solver *S, double *ux, double *uy, double *uz) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || !ux || !uy || !uz) return fail(X3D_ERR_ARG, "null pointer");
  double *dst[3] = {ux, uy, uz};
  X3D_TRY(sync_velocity(*c, S));
  for (int i = 0; i < 3; ++i)
    X3D_HIP(hipMemcpyAsync(dst[i], S->u64[i], S->n * 8, hipMemcpyDeviceToDevice, c->stream));
  X3D_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

int x3d_solver_step(x3d_solver *S, int nsteps) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || nsteps < 0) return fail(X3D_ERR_ARG, "bad arguments");
  for (int k = 0; k < nsteps; ++k) {
    if (c->p.itimescheme == 5) {
      if (S->precision == 0) X3D_TRY(step_rk3<double>(*c, S));
      else X3D_TRY(step_rk3<float>(*c, S));
    } else if (S->precision == 0) {
      X3D_TRY(step_once<double>(*c, S));
    } else {
      X3D_TRY(step_once<float>(*c, S));
    }
    S->have_pressure = true;
  }
  return X3D_OK;
}

int x3d_solver_tgv_stats(x3d_solver *S, double out4[4]) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || !out4) return fail(X3D_ERR_ARG, "null pointer");
  if (S->precision == 0) return tgv_stats<double>(*c, S, out4);
  return tgv_stats<float>(*c, S, out4);
}

int x3d_solver_divergence_stats(x3d_solver *S, double out2[2]) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || !out2) return fail(X3D_ERR_ARG, "null pointer");
  // FP64 divergence of the FP64 velocity, whatever the compute precision
  const size_t np = (size_t)c->nxm * c->nym * c->nzm;
  const size_t nw = S->n > np ? S->n : np;
  for (int i = 0; i < 4; ++i)
    if (!S->w64[i]) X3D_HIP(hipMalloc(&S->w64[i], nw * 8));
  double *W[4] = {S->w64[0], S->w64[1], S->w64[2], S->w64[3]};
  X3D_TRY(sync_velocity(*c, S));
  const double *u[3] = {S->u64[0], S->u64[1], S->u64[2]};
  X3D_TRY(divergence<double>(*c, u, W, S->nx, S->ny, S->nz));
  X3D_TRY(launch_max_mean<double>(W[0], np, S->partial, S->red, c->stream));
  X3D_HIP(hipMemcpyAsync(out2, S->red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  X3D_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

int x3d_solver_divu_stats(x3d_solver *S, double out2[2]) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || !out2) return fail(X3D_ERR_ARG, "null pointer");
  if (S->divp_n == 0) return fail(X3D_ERR_STATE, "div u* statistics: no step has run since the field was set");
  const size_t np = (size_t)c->nxm * c->nym * c->nzm;
  X3D_TRY(launch_finish_absmax(S->divp, S->divp_n, np, S->red, c->stream));
  X3D_HIP(hipMemcpyAsync(out2, S->red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  X3D_HIP(hipStreamSynchronize(c->stream));
  // tmax starts at -1609 in the reference
  if (out2[0] < -1609.0) out2[0] = -1609.0;
  return X3D_OK;
}

int x3d_solver_set_les(x3d_solver *S, int model, double cs, int iconserv, double beta, x3d_mlp *mlp, double gscale,
                       double delta0) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S) return fail(X3D_ERR_ARG, "null solver");
  if (model < 0 || model > 3)
    return fail(X3D_ERR_ARG, "LES model must be 0 (none), 1 (Smagorinsky), 2 or 3 (learned)");
  if (iconserv != 0 && iconserv != 1) return fail(X3D_ERR_ARG, "iconserv must be 0 or 1");
  if (model >= 2 && (!mlp || !(gscale > 0) || !(delta0 > 0)))
    return fail(X3D_ERR_ARG, "learned closure needs a network and positive gscale, delta0");
  if (!(cs >= 0)) return fail(X3D_ERR_ARG, "cs must be >= 0");
  if (model && c->p.istret != 0)
    return fail(X3D_ERR_UNSUPPORTED, "LES on a stretched y mesh (the filter width varies with y) is not provided");
  if (model && !S->lesw) {
    const size_t es = S->precision == 0 ? 8 : 4;
    X3D_HIP(hipMalloc(&S->lesw, 20 * S->n * es));
    X3D_HIP(hipMalloc(&S->nfb, sizeof(int)));
    X3D_HIP(hipMemset(S->nfb, 0, sizeof(int)));
  }
  S->les = model;
  S->cs = cs;
  S->iconserv = iconserv;
  S->beta = beta;
  S->mlp = model >= 2 ? reinterpret_cast<MlpDev *>(mlp) : nullptr;
  S->gscale = gscale;
  S->delta0 = delta0;
  return X3D_OK;
}

int x3d_solver_sgs(x3d_solver *S, double *sx, double *sy, double *sz) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || !sx || !sy || !sz) return fail(X3D_ERR_ARG, "null pointer");
  if (!S->les) return fail(X3D_ERR_STATE, "no LES model set");
  double *out[3] = {sx, sy, sz};
  if (S->precision == 0) {
    const double *u[3] = {S->u64[0], S->u64[1], S->u64[2]};
    X3D_TRY(compute_sgs<double>(*c, S, u, out));
  } else {
    const float *u[3] = {S->u32[0], S->u32[1], S->u32[2]};
    float *sg[3];
    for (int i = 0; i < 3; ++i) sg[i] = (float *)S->lesw + (17 + i) * S->n;
    X3D_TRY(compute_sgs<float>(*c, S, u, sg));
    for (int i = 0; i < 3; ++i) X3D_TRY(launch_widen(sg[i], out[i], S->n, c->stream));
  }
  X3D_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

int x3d_solver_les_fallbacks(x3d_solver *S, long long *count) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || !count) return fail(X3D_ERR_ARG, "null pointer");
  int h = 0;
  if (S->nfb) {
    X3D_HIP(hipMemcpyAsync(&h, S->nfb, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    X3D_HIP(hipStreamSynchronize(c->stream));
    X3D_HIP(hipMemsetAsync(S->nfb, 0, sizeof(int), c->stream));
    X3D_HIP(hipStreamSynchronize(c->stream));
  }
  *count = h;
  return X3D_OK;
}

int x3d_solver_momentum_yz(x3d_solver *S, int nrep, double *bytes_per_launch) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || nrep < 0 || !bytes_per_launch) return fail(X3D_ERR_ARG, "bad arguments");
  if (!S->fused) return fail(X3D_ERR_UNSUPPORTED, "momentum sweeps: the solver does not use the fused path");
  if (S->precision == 0) return momentum_yz<double>(*c, S, nrep, bytes_per_launch);
  return momentum_yz<float>(*c, S, nrep, bytes_per_launch);
}

int x3d_solver_init_channel(x3d_solver *S, double init_noise, unsigned long long seed) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S) return fail(X3D_ERR_ARG, "null solver");
  if (!(init_noise >= 0)) return fail(X3D_ERR_ARG, "init_noise must be >= 0");
  if (seed >= (1ull << 24)) return fail(X3D_ERR_ARG, "seed must be below 2^24");
  // y - yly/2 of the velocity points (Case-Channel.f90:100-102)
  std::vector<double> yc(S->ny);
  for (int j = 0; j < S->ny; ++j)
    yc[j] = c->p.istret != 0 ? c->yp[j] - c->p.yly * 0.5 : (double)j * c->dir[1].d - c->p.yly * 0.5;
  if (!S->ycoord) X3D_HIP(hipMalloc(&S->ycoord, S->ny * sizeof(double)));
  X3D_HIP(hipMemcpyAsync(S->ycoord, yc.data(), S->ny * sizeof(double), hipMemcpyHostToDevice, c->stream));
  X3D_TRY(launch_channel_init(S->u64[0], S->u64[1], S->u64[2], S->ycoord, S->nx, S->ny, S->nz, init_noise, seed,
                              c->stream));
  return reset_history(*c, S);
}

int x3d_solver_set_forcing(x3d_solver *S, int mode, double value) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S) return fail(X3D_ERR_ARG, "null solver");
  if (mode < 0 || mode > 2) return fail(X3D_ERR_ARG, "forcing must be 0 (none), 1 (flow rate) or 2 (pressure gradient)");
  if (mode == 1 && (c->dir[1].bc1 == 0 || c->dir[1].bcn == 0))
    return fail(X3D_ERR_UNSUPPORTED, "the constant flow rate is defined for a y-bounded channel");
  S->forcing = mode;
  S->fvalue = value;
  return X3D_OK;
}

int x3d_solver_set_rotation(x3d_solver *S, double wrotation, int spinup_time) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S) return fail(X3D_ERR_ARG, "null solver");
  if (!std::isfinite(wrotation) || spinup_time < 0) return fail(X3D_ERR_ARG, "bad rotation arguments");
  if (S->fused && spinup_time > 0)
    return fail(X3D_ERR_UNSUPPORTED, "the spin-up rotation runs on the per-operator step (bounded y, or RK3)");
  S->wrotation = wrotation;
  S->spinup_time = spinup_time;
  return X3D_OK;
}

int x3d_solver_bulk_velocity(x3d_solver *S, double *ub) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || !ub) return fail(X3D_ERR_ARG, "null pointer");
  const double coeff = c->dir[1].d / (c->p.yly * (double)((long long)S->nx * S->nz));
  X3D_TRY(sync_velocity(*c, S));
  X3D_TRY(launch_bulk(S->u64[0], c->p.istret != 0 ? c->ppy64 : nullptr, S->nx, S->ny, S->nz, coeff, 0.0, S->partial,
                      S->red, c->stream));
  X3D_HIP(hipMemcpyAsync(ub, S->red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  X3D_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

int x3d_solver_itime(x3d_solver *S, int *itime) {
  if (!S || !itime) return fail(X3D_ERR_ARG, "null pointer");
  *itime = S->itime;
  return X3D_OK;
}

// ---------------------------------------------------------------- checkpoint / restart
// restart() of tools.f90:205-514 for this solver's state: the checkpoint
// directory <rundir>/checkpoint holds one file per variable the reference
// writes (ux, uy, uz; dux-2 .. duz-2 and dux-3 .. duz-3 for AB3; pp on the
// pressure grid), each the global array in Fortran order as native FP64
// (decomp_2d_write_one with reduce_prec=.false.), and <rundir>/restart.info
// the &Time / &NumParam namelists with the reference's formats
// (tools.f90:343-373).  Reading restores the velocity, the AB3 history, the
// pressure, itime (the &Time namelist) and, with y walls, the wall planes of
// the pressure gradient that gradp reconstructs from pp (tools.f90:493-500).
namespace {

int write_raw(const std::string &path, const double *a, size_t n) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return fail(X3D_ERR_IO, "checkpoint: cannot create " + path + ": " + std::strerror(errno));
  const size_t w = std::fwrite(a, sizeof(double), n, f);
  const int cl = std::fclose(f);
  if (w != n || cl != 0) return fail(X3D_ERR_IO, "checkpoint: short write to " + path);
  return X3D_OK;
}

int read_raw(const std::string &path, double *a, size_t n) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return fail(X3D_ERR_IO, "restart: cannot open " + path + ": " + std::strerror(errno));
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz != (long)(n * sizeof(double))) {
    std::fclose(f);
    return fail(X3D_ERR_IO, "restart: " + path + " holds " + std::to_string(sz) + " bytes, the grid needs " +
                                std::to_string(n * sizeof(double)));
  }
  const size_t r = std::fread(a, sizeof(double), n, f);
  std::fclose(f);
  if (r != n) return fail(X3D_ERR_IO, "restart: short read from " + path);
  return X3D_OK;
}

// device field of compute precision es (8 or 4) <-> FP64 host buffer
int field_to_host(const void *dev, size_t es, size_t n, std::vector<double> &h, hipStream_t st) {
  h.resize(n);
  if (es == 8) {
    X3D_HIP(hipMemcpyAsync(h.data(), dev, n * 8, hipMemcpyDeviceToHost, st));
    X3D_HIP(hipStreamSynchronize(st));
  } else {
    std::vector<float> t(n);
    X3D_HIP(hipMemcpyAsync(t.data(), dev, n * 4, hipMemcpyDeviceToHost, st));
    X3D_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < n; ++i) h[i] = t[i];
  }
  return X3D_OK;
}

int field_from_host(void *dev, size_t es, const std::vector<double> &h, hipStream_t st) {
  const size_t n = h.size();
  if (es == 8) {
    X3D_HIP(hipMemcpyAsync(dev, h.data(), n * 8, hipMemcpyHostToDevice, st));
    X3D_HIP(hipStreamSynchronize(st));
  } else {
    std::vector<float> t(n);
    for (size_t i = 0; i < n; ++i) t[i] = (float)h[i];
    X3D_HIP(hipMemcpyAsync(dev, t.data(), n * 4, hipMemcpyHostToDevice, st));
    X3D_HIP(hipStreamSynchronize(st));
  }
  return X3D_OK;
}

const char *const VNAME[3] = {"ux", "uy", "uz"};
const char *const DNAME[3] = {"dux", "duy", "duz"};

}  // namespace

int x3d_solver_checkpoint(x3d_solver *S, const char *rundir, double t) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || !rundir) return fail(X3D_ERR_ARG, "null pointer");
  if (!S->have_pressure) return fail(X3D_ERR_STATE, "checkpoint: no step has run since the field was set");
  const std::string dir = std::string(rundir) + "/checkpoint";
  if (mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST)
    return fail(X3D_ERR_IO, "checkpoint: cannot create " + dir + ": " + std::strerror(errno));
  const size_t es = S->precision == 0 ? 8 : 4;
  std::vector<double> h;
  X3D_TRY(sync_velocity(*c, S));
  X3D_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < 3; ++i) {
    X3D_TRY(field_to_host(S->u64[i], 8, S->n, h, c->stream));
    X3D_TRY(write_raw(dir + "/" + VNAME[i], h.data(), S->n));
  }
  if (c->p.itimescheme == 3) {
    // after a step f(n) sits in slot[1] and f(n-1) in slot[2]: dux1(:,:,:,2)
    // and dux1(:,:,:,3) of time_integrators.f90:85-100
    for (int lag = 2; lag <= 3; ++lag)
      for (int i = 0; i < 3; ++i) {
        X3D_TRY(field_to_host(S->hist[i][S->slot[lag - 1]], es, S->n, h, c->stream));
        X3D_TRY(write_raw(dir + "/" + DNAME[i] + "-" + std::to_string(lag), h.data(), S->n));
      }
  }
  const size_t np = (size_t)c->nxm * c->nym * c->nzm;
  X3D_TRY(field_to_host(S->w[0], es, np, h, c->stream));
  X3D_TRY(write_raw(dir + "/pp", h.data(), np));
  const std::string info = std::string(rundir) + "/restart.info";
  FILE *f = std::fopen(info.c_str(), "w");
  if (!f) return fail(X3D_ERR_IO, "checkpoint: cannot create " + info + ": " + std::strerror(errno));
  const char *bar = "!=========================";
  std::fprintf(f, "%s\n&Time\n%s\n", bar, bar);
  std::fprintf(f, "tfield=   %16.4f\n", t);
  std::fprintf(f, "itime=    %16d\n", S->itime);
  std::fprintf(f, "/End\n%s\n&NumParam\n%s\n", bar, bar);
  std::fprintf(f, "nx=       %16d\nny=       %16d\nnz=       %16d\n", S->nx, S->ny, S->nz);
  std::fprintf(f, "Lx=       %16.4f\nLy=       %16.4f\nLz=       %16.4f\n", c->p.xlx, c->p.yly, c->p.zlz);
  std::fprintf(f, "istret=   %16d\n", c->p.istret);
  std::fprintf(f, "beta=     %16.12f\n", c->p.beta);
  std::fprintf(f, "iscalar=  %16d\nnumscalar=%16d\n", 0, 0);
  std::fprintf(f, "itimescheme=%14d\n", c->p.itimescheme);
  std::fprintf(f, "iimplicit=%16d\n", 0);
  std::fprintf(f, "/End\n%s\n", bar);
  if (std::fclose(f) != 0) return fail(X3D_ERR_IO, "checkpoint: cannot write " + info);
  return X3D_OK;
}

int x3d_solver_restart(x3d_solver *S, const char *rundir, double *t) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!S || !rundir) return fail(X3D_ERR_ARG, "null pointer");
  const std::string dir = std::string(rundir) + "/checkpoint";
  const size_t es = S->precision == 0 ? 8 : 4;
  // &Time namelist: tfield and itime (tools.f90:457-470)
  const std::string info = std::string(rundir) + "/restart.info";
  double tfield = 0.0;
  int itime = -1;
  if (FILE *f = std::fopen(info.c_str(), "r")) {
    char line[256];
    bool in_time = false;
    while (std::fgets(line, sizeof line, f)) {
      const char *q = line;
      while (*q == ' ') ++q;
      if (!std::strncmp(q, "&Time", 5)) in_time = true;
      else if (!std::strncmp(q, "/End", 4)) in_time = false;
      else if (in_time && !std::strncmp(q, "tfield=", 7)) tfield = std::strtod(q + 7, nullptr);
      else if (in_time && !std::strncmp(q, "itime=", 6)) itime = (int)std::strtol(q + 6, nullptr, 10);
    }
    std::fclose(f);
  } else {
    return fail(X3D_ERR_IO, "restart: cannot open " + info + ": " + std::strerror(errno));
  }
  if (itime < 0) return fail(X3D_ERR_IO, "restart: " + info + " has no itime in its &Time namelist");
  std::vector<double> h(S->n);
  X3D_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < 3; ++i) {
    X3D_TRY(read_raw(dir + "/" + VNAME[i], h.data(), S->n));
    X3D_TRY(field_from_host(S->u64[i], 8, h, c->stream));
  }
  X3D_TRY(narrow_velocity(*c, S));
  S->slot[0] = 0;
  S->slot[1] = 1;
  S->slot[2] = 2;
  if (c->p.itimescheme == 3)
    for (int lag = 2; lag <= 3; ++lag)
      for (int i = 0; i < 3; ++i) {
        X3D_TRY(read_raw(dir + "/" + DNAME[i] + "-" + std::to_string(lag), h.data(), S->n));
        X3D_TRY(field_from_host(S->hist[i][lag - 1], es, h, c->stream));
      }
  const size_t np = (size_t)c->nxm * c->nym * c->nzm;
  std::vector<double> hp(np);
  X3D_TRY(read_raw(dir + "/pp", hp.data(), np));
  X3D_TRY(field_from_host(S->w[0], es, hp, c->stream));
  S->itime = itime;
  S->have_pressure = true;
  if (S->wall1 || S->walln) {
    // gradp on the restored pressure stores the wall planes over gdt of the
    // last sub-step (itr = 3 for RK3, 1 for AB3)
    const double dt = c->p.dt;
    const double gdt = c->p.itimescheme == 5 ? (3.0 / 4.0) * dt + (-5.0 / 12.0) * dt
                                             : (23.0 / 12.0) * dt + (-(16.0 / 12.0) * dt) + (5.0 / 12.0) * dt;
    if (S->precision == 0) {
      double *u[3] = {S->u64[0], S->u64[1], S->u64[2]};
      double *W[5] = {(double *)S->w[0], (double *)S->w[1], (double *)S->w[2], (double *)S->w[3], (double *)S->w[4]};
      X3D_TRY(gradp_correct<double>(*c, S, u, W, gdt, false));
    } else {
      float *u[3] = {S->u32[0], S->u32[1], S->u32[2]};
      float *W[5] = {(float *)S->w[0], (float *)S->w[1], (float *)S->w[2], (float *)S->w[3], (float *)S->w[4]};
      X3D_TRY(gradp_correct<float>(*c, S, u, W, gdt, false));
    }
  }
  X3D_HIP(hipStreamSynchronize(c->stream));
  if (t) *t = tfield;
  return X3D_OK;
}

}  // extern "C"
