// Host FP64 computation of every coefficient set of schemes.f90 and upload of
// FP64/FP32 device copies.  The arithmetic follows schemes.f90 statement by
// statement so that the arrays equal the reference's module arrays.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "x3d_internal.h"

namespace x3d {

namespace {

// prepare(b,c,f,s,w,n), schemes.f90:413-439
void prepare(const std::vector<double> &b, const std::vector<double> &c, const std::vector<double> &f,
             std::vector<double> &s, std::vector<double> &w) {
  const int n = (int)c.size();
  w = c;
  s.assign(n, 0.0);
  for (int i = 1; i < n; ++i) {
    s[i] = b[i - 1] / w[i - 1];
    w[i] = w[i] - f[i - 1] * s[i];
  }
  for (int i = 0; i < n; ++i) w[i] = 1.0 / w[i];
}

struct D1 {
  double alfai, afi, bfi, alfa1, af1, bf1, cf1, alfa2, af2, alfan, afn, bfn, cfn, alfam, afm;
  std::vector<double> ff, fs, fw, ffp, fsp, fwp;
};

// first_derivative, schemes.f90:443-599
int first_derivative(double d, int n, int ncl1, int ncln, int ifirstder, D1 &o) {
  if (ifirstder == 1) {
    o.alfai = 0.0; o.afi = 1.0 / (2.0 * d); o.bfi = 0.0;
    o.alfa1 = o.af1 = o.bf1 = o.cf1 = o.alfa2 = o.af2 = 0.0;
    o.alfam = o.afm = o.alfan = o.afn = o.bfn = o.cfn = 0.0;
  } else if (ifirstder == 4) {
    o.alfai = 1.0 / 3.0; o.afi = (7.0 / 9.0) / d; o.bfi = (1.0 / 36.0) / d;
    o.alfa1 = 2.0; o.af1 = -(5.0 / 2.0) / d; o.bf1 = 2.0 / d; o.cf1 = 0.5 / d;
    o.alfa2 = 1.0 / 4.0; o.af2 = (3.0 / 4.0) / d;
    o.alfan = 2.0; o.afn = -(5.0 / 2.0) / d; o.bfn = 2.0 / d; o.cfn = 0.5 / d;
    o.alfam = 1.0 / 4.0; o.afm = (3.0 / 4.0) / d;
  } else {
    return fail(X3D_ERR_UNSUPPORTED, "ifirstder must be 1 or 4 (schemes.f90:467-486)");
  }
  std::vector<double> ff(n, 0.0), fb(n, 0.0), fc(n, 0.0);
  const double al = o.alfai;
  if (ncl1 == 0) {
    ff[0] = al; ff[1] = al; fc[0] = 2.0; fc[1] = 1.0; fb[0] = al; fb[1] = al;
  } else if (ncl1 == 1) {
    ff[0] = al + al; ff[1] = al; fc[0] = 1.0; fc[1] = 1.0; fb[0] = al; fb[1] = al;
  } else {
    ff[0] = o.alfa1; ff[1] = o.alfa2; fc[0] = 1.0; fc[1] = 1.0; fb[0] = o.alfa2; fb[1] = al;
  }
  if (ncln == 0) {
    ff[n - 3] = al; ff[n - 2] = al; ff[n - 1] = 0.0;
    fc[n - 3] = 1.0; fc[n - 2] = 1.0; fc[n - 1] = 1.0 + al * al;
    fb[n - 3] = al; fb[n - 2] = al; fb[n - 1] = 0.0;
  } else if (ncln == 1) {
    ff[n - 3] = al; ff[n - 2] = al; ff[n - 1] = 0.0;
    fc[n - 3] = 1.0; fc[n - 2] = 1.0; fc[n - 1] = 1.0;
    fb[n - 3] = al; fb[n - 2] = al + al; fb[n - 1] = 0.0;
  } else {
    ff[n - 3] = al; ff[n - 2] = o.alfam; ff[n - 1] = 0.0;
    fc[n - 3] = 1.0; fc[n - 2] = 1.0; fc[n - 1] = 1.0;
    fb[n - 3] = o.alfam; fb[n - 2] = o.alfan; fb[n - 1] = 0.0;
  }
  for (int i = 2; i < n - 3; ++i) { ff[i] = al; fc[i] = 1.0; fb[i] = al; }
  std::vector<double> ffp = ff;
  prepare(fb, fc, ff, o.fs, o.fw);
  if (ncl1 == 1) ffp[0] = 0.0;
  if (ncln == 1) fb[n - 2] = 0.0;
  prepare(fb, fc, ffp, o.fsp, o.fwp);
  o.ff = ff;
  o.ffp = ffp;
  return X3D_OK;
}

struct D2 {
  double alsai, asi, bsi, csi, dsi;
  double alsa1, as1, bs1, cs1, ds1, alsa2, as2, alsa3, as3, bs3, alsa4, as4, bs4, cs4;
  double alsan, asn, bsn, csn, dsn, alsam, asm_, alsat, ast, bst, alsatt, astt, bstt, cstt;
  std::vector<double> sf, ss, sw, sfp, ssp, swp;
};

// second_derivative, schemes.f90:602-856
int second_derivative(double d2, int n, int ncl1, int ncln, int isecondder, double nu0nu, double cnu,
                      D2 &o) {
  const double pi = std::acos(-1.0);
  if (isecondder == 1) {
    o.alsai = 0.0; o.asi = 1.0 / d2; o.bsi = 0.0; o.csi = 0.0; o.dsi = 0.0;
  } else if (isecondder == 4) {
    o.alsai = 2.0 / 11.0; o.asi = (12.0 / 11.0) / d2; o.bsi = (3.0 / 44.0) / d2; o.csi = 0.0; o.dsi = 0.0;
  } else if (isecondder == 5) {
    const double dpis3 = 2.0 * pi / 3.0;
    const double kppkc = pi * pi * (1.0 + nu0nu);
    const double kppkm = dpis3 * dpis3 * (1.0 + cnu * nu0nu);
    const double xnpi2 = kppkc, xmpi2 = kppkm;
    const double den = 405.0 * xnpi2 - 640.0 * xmpi2 + 144.0;
    o.alsai = 0.5 - (320.0 * xmpi2 - 1296.0) / den;
    o.asi = -(4329.0 * xnpi2 / 8.0 - 32.0 * xmpi2 - 140.0 * xnpi2 * xmpi2 + 286.0) / den / d2;
    o.bsi = (2115.0 * xnpi2 - 1792.0 * xmpi2 - 280.0 * xnpi2 * xmpi2 + 1328.0) / den / (4.0 * d2);
    o.csi = -(7695.0 * xnpi2 / 8.0 + 288.0 * xmpi2 - 180.0 * xnpi2 * xmpi2 - 2574.0) / den / (9.0 * d2);
    o.dsi = (198.0 * xnpi2 + 128.0 * xmpi2 - 40.0 * xnpi2 * xmpi2 - 736.0) / den / (16.0 * d2);
  } else {
    return fail(X3D_ERR_UNSUPPORTED, "isecondder must be 1, 4 or 5 (schemes.f90:633-695)");
  }
  o.alsa1 = 11.0; o.as1 = 13.0 / d2; o.bs1 = -27.0 / d2; o.cs1 = 15.0 / d2; o.ds1 = -1.0 / d2;
  if (isecondder == 1) { o.alsa2 = 0.0; o.as2 = 1.0 / d2; o.alsam = 0.0; o.asm_ = 1.0 / d2; }
  else { o.alsa2 = 0.1; o.as2 = (6.0 / 5.0) / d2; o.alsam = 0.1; o.asm_ = (6.0 / 5.0) / d2; }
  o.alsa3 = 2.0 / 11.0; o.as3 = (12.0 / 11.0) / d2; o.bs3 = (3.0 / 44.0) / d2;
  o.alsa4 = 2.0 / 11.0; o.as4 = (12.0 / 11.0) / d2; o.bs4 = (3.0 / 44.0) / d2; o.cs4 = 0.0;
  o.alsan = 11.0; o.asn = 13.0 / d2; o.bsn = -27.0 / d2; o.csn = 15.0 / d2; o.dsn = -1.0 / d2;
  o.alsat = 2.0 / 11.0; o.ast = (12.0 / 11.0) / d2; o.bst = (3.0 / 44.0) / d2;
  o.alsatt = 2.0 / 11.0; o.astt = (12.0 / 11.0) / d2; o.bstt = (3.0 / 44.0) / d2; o.cstt = 0.0;
  std::vector<double> sf(n, 0.0), sb(n, 0.0), sc(n, 0.0);
  const double al = o.alsai;
  if (ncl1 == 0) {
    for (int i = 0; i < 4; ++i) { sf[i] = al; sc[i] = 1.0; sb[i] = al; }
    sc[0] = 2.0;
  } else if (ncl1 == 1) {
    for (int i = 0; i < 4; ++i) { sf[i] = al; sc[i] = 1.0; sb[i] = al; }
    sf[0] = al + al;
  } else {
    sf[0] = o.alsa1; sf[1] = o.alsa2; sf[2] = o.alsa3; sf[3] = o.alsa4;
    for (int i = 0; i < 4; ++i) sc[i] = 1.0;
    sb[0] = o.alsa2; sb[1] = o.alsa3; sb[2] = o.alsa4; sb[3] = al;
  }
  if (ncln == 0) {
    for (int i = n - 5; i < n - 1; ++i) { sf[i] = al; sc[i] = 1.0; sb[i] = al; }
    sf[n - 1] = 0.0; sc[n - 1] = 1.0 + al * al; sb[n - 1] = 0.0;
  } else if (ncln == 1) {
    for (int i = n - 5; i < n - 1; ++i) { sf[i] = al; sc[i] = 1.0; sb[i] = al; }
    sf[n - 1] = 0.0; sc[n - 1] = 1.0; sb[n - 2] = al + al; sb[n - 1] = 0.0;
  } else {
    sf[n - 5] = al; sf[n - 4] = o.alsatt; sf[n - 3] = o.alsat; sf[n - 2] = o.alsam; sf[n - 1] = 0.0;
    for (int i = n - 5; i < n; ++i) sc[i] = 1.0;
    sb[n - 5] = o.alsatt; sb[n - 4] = o.alsat; sb[n - 3] = o.alsam; sb[n - 2] = o.alsan; sb[n - 1] = 0.0;
  }
  for (int i = 4; i < n - 5; ++i) { sf[i] = al; sc[i] = 1.0; sb[i] = al; }
  std::vector<double> sfp = sf;
  if (ncl1 == 1) sf[0] = 0.0;
  prepare(sb, sc, sf, o.ss, o.sw);
  prepare(sb, sc, sfp, o.ssp, o.swp);
  if (ncln == 1) {
    sb[n - 2] = 0.0;
    prepare(sb, sc, sf, o.ss, o.sw);
  }
  o.sf = sf;
  o.sfp = sfp;
  return X3D_OK;
}

struct ST {
  double alcai, aci, bci, ailcai, aici, bici, cici, dici;
  std::vector<double> cfx6, csx6, cwx6, cfxp6, csxp6, cwxp6, cifx6, cisx6, ciwx6, cifxp6, cisxp6, ciwxp6;
  std::vector<double> cfi6, csi6, cwi6, cfip6, csip6, cwip6, cifi6, cisi6, ciwi6, cifip6, cisip6, ciwip6;
};

// interpolation, schemes.f90:860-1066
int interpolation(double dx, int nxm, int nx, int ncl1, int ncln, int ifirstder, int ipinter, ST &o) {
  if (ifirstder == 1) { o.alcai = 0.0; o.aci = 1.0 / dx; o.bci = 0.0; }
  else { o.alcai = 9.0 / 62.0; o.aci = (63.0 / 62.0) / dx; o.bci = (17.0 / 62.0) / 3.0 / dx; }
  if (ifirstder == 1) {
    o.ailcai = 0.0; o.aici = 0.5; o.bici = 0.0; o.cici = 0.0; o.dici = 0.0;
  } else if (ipinter == 1) {
    o.ailcai = 3.0 / 10.0; o.aici = 3.0 / 4.0; o.bici = 1.0 / (2.0 * 10.0); o.cici = 0.0; o.dici = 0.0;
  } else if (ipinter == 2) {
    o.ailcai = 0.461658;
    o.dici = 0.00293016;
    o.aici = 1.0 / 64.0 * (75.0 + 70.0 * o.ailcai - 320.0 * o.dici);
    o.bici = 1.0 / 128.0 * (126.0 * o.ailcai - 25.0 + 1152.0 * o.dici);
    o.cici = 1.0 / 128.0 * (-10.0 * o.ailcai + 3.0 - 640.0 * o.dici);
    o.aici /= 2.0; o.bici /= 2.0; o.cici /= 2.0; o.dici /= 2.0;
  } else if (ipinter == 3) {
    o.ailcai = 0.49;
    o.aici = 1.0 / 128.0 * (75.0 + 70.0 * o.ailcai);
    o.bici = 1.0 / 256.0 * (126.0 * o.ailcai - 25.0);
    o.cici = 1.0 / 256.0 * (-10.0 * o.ailcai + 3.0);
    o.dici = 0.0;
  } else {
    return fail(X3D_ERR_UNSUPPORTED, "ipinter must be 1, 2 or 3 (schemes.f90:907-931)");
  }
  auto tri_m = [&](double al, std::vector<double> &f, std::vector<double> &c, std::vector<double> &b) {
    f.assign(nxm, 0.0); c.assign(nxm, 0.0); b.assign(nxm, 0.0);
    f[0] = al; f[1] = al; f[nxm - 3] = al; f[nxm - 2] = al; f[nxm - 1] = 0.0;
    c[0] = ncl1 == 0 ? 2.0 : 1.0 + al;
    c[1] = 1.0; c[nxm - 3] = 1.0; c[nxm - 2] = 1.0;
    c[nxm - 1] = ncln == 0 ? 1.0 + al * al : 1.0 + al;
    b[0] = al; b[1] = al; b[nxm - 3] = al; b[nxm - 2] = al; b[nxm - 1] = 0.0;
    for (int i = 2; i < nxm - 3; ++i) { f[i] = al; c[i] = 1.0; b[i] = al; }
  };
  auto tri_n = [&](double al, std::vector<double> &f, std::vector<double> &c, std::vector<double> &b) {
    f.assign(nx, 0.0); c.assign(nx, 0.0); b.assign(nx, 0.0);
    f[0] = al + al; f[1] = al; f[nx - 3] = al; f[nx - 2] = al; f[nx - 1] = 0.0;
    c[0] = 1.0; c[1] = 1.0; c[nx - 3] = 1.0; c[nx - 2] = 1.0; c[nx - 1] = 1.0;
    b[0] = al; b[1] = al; b[nx - 3] = al; b[nx - 2] = al + al; b[nx - 1] = 0.0;
    for (int i = 2; i < nx - 3; ++i) { f[i] = al; c[i] = 1.0; b[i] = al; }
  };
  std::vector<double> ccx6, cbx6, cci6, cbi6, cicx6, cibx6, cici6, cibi6;
  tri_m(o.alcai, o.cfx6, ccx6, cbx6);
  tri_n(o.alcai, o.cfi6, cci6, cbi6);
  tri_m(o.ailcai, o.cifx6, cicx6, cibx6);
  tri_n(o.ailcai, o.cifi6, cici6, cibi6);
  o.cfxp6 = o.cfx6; o.cifxp6 = o.cifx6; o.cifip6 = o.cifi6; o.cfip6 = o.cfi6;
  o.cfxp6[0] = 0.0;
  o.cfip6[0] = 0.0;
  prepare(cbx6, ccx6, o.cfx6, o.csx6, o.cwx6);
  prepare(cbx6, ccx6, o.cfxp6, o.csxp6, o.cwxp6);
  prepare(cibx6, cicx6, o.cifx6, o.cisx6, o.ciwx6);
  prepare(cibx6, cicx6, o.cifxp6, o.cisxp6, o.ciwxp6);
  prepare(cbi6, cci6, o.cfi6, o.csi6, o.cwi6);
  prepare(cbi6, cci6, o.cfip6, o.csip6, o.cwip6);
  prepare(cibi6, cici6, o.cifi6, o.cisi6, o.ciwi6);
  prepare(cibi6, cici6, o.cifip6, o.cisip6, o.ciwip6);
  if (ncln == 1 || ncln == 2) {
    cbx6[nxm - 2] = 0.0;
    cibx6[nxm - 1] = 0.0;
    cbi6[nx - 2] = 0.0;
    cibi6[nx - 1] = 0.0;
    prepare(cbx6, ccx6, o.cfxp6, o.csxp6, o.cwxp6);
    prepare(cibx6, cicx6, o.cifxp6, o.cisxp6, o.ciwxp6);
    prepare(cbi6, cci6, o.cfip6, o.csip6, o.cwip6);
    prepare(cibi6, cici6, o.cifip6, o.cisip6, o.ciwip6);
  }
  return X3D_OK;
}

// Sherman-Morrison vector of the periodic solve: the data-independent r of
// derive.f90:22-58 (r = (-1,0,...,0,alfa) through the same recurrences).
void sm_vector(const std::vector<double> &ff, const std::vector<double> &fs, const std::vector<double> &fw,
               double alfa, std::vector<double> &r, double &den) {
  const int n = (int)ff.size();
  r.assign(n, 0.0);
  r[0] = -1.0;
  r[n - 1] = alfa;
  for (int i = 1; i < n; ++i) r[i] = r[i] - r[i - 1] * fs[i];
  r[n - 1] = r[n - 1] * fw[n - 1];
  for (int i = n - 2; i >= 0; --i) r[i] = (r[i] - ff[i] * r[i + 1]) * fw[i];
  den = 1.0 + r[0] - alfa * r[n - 1];
}

int add_set(Context &c, const std::string &name, const std::vector<double> &ff, const std::vector<double> &fs,
            const std::vector<double> &fw, int periodic, double alfa) {
  TriSet s;
  s.name = name;
  s.ff = ff; s.fs = fs; s.fw = fw;
  s.periodic = periodic;
  s.alfa = alfa;
  if (periodic) sm_vector(ff, fs, fw, alfa, s.sm, s.sm_den);
  else s.sm.assign(ff.size(), 0.0);
  c.sets.push_back(std::move(s));
  return (int)c.sets.size() - 1;
}

template <typename T>
void cast_to(const Sten<double> &a, Sten<T> &b) {
  const double *src = reinterpret_cast<const double *>(&a);
  T *dst = reinterpret_cast<T *>(&b);
  static_assert(sizeof(Sten<double>) % sizeof(double) == 0, "Sten layout");
  for (size_t i = 0; i < sizeof(Sten<double>) / sizeof(double); ++i) dst[i] = (T)src[i];
}

}  // namespace

TriDev<double> TriSet::dev64() const {
  TriDev<double> d;
  d.pk = d64;
  d.alfa = alfa; d.sm_den = sm_den; d.rden = 1.0 / sm_den;
  d.n = (int)ff.size(); d.periodic = periodic; d.kdepth = kd64;
  return d;
}

TriDev<float> TriSet::dev32() const {
  TriDev<float> d;
  d.pk = d32;
  d.alfa = (float)alfa; d.sm_den = (float)sm_den; d.rden = (float)(1.0 / sm_den);
  d.n = (int)ff.size(); d.periodic = periodic; d.kdepth = kd32;
  return d;
}

bool same_set(const TriSet &a, const TriSet &b) {
  return &a == &b || (a.ff == b.ff && a.fs == b.fs && a.fw == b.fw && a.sm == b.sm && a.alfa == b.alfa &&
                      a.sm_den == b.sm_den);
}

// Smallest K such that the elimination factors of any K consecutive segments
// of L points multiply to less than tol*(1-rho) in magnitude, rho being the
// largest single-segment product: the carries of segments further back then
// change no result by more than tol times the largest carry.  Forward factors
// are fs, backward factors ff*fw (the recurrences of prepare, schemes.f90:
// 413-439).  Returns P (the full chain) when no shorter depth qualifies.
int chain_depth(const TriSet &s, int L, double tol) {
  const int n = (int)s.fs.size();
  const int P = (n + L - 1) / L;
  std::vector<double> gfw(P, 1.0), gbw(P, 1.0);
  for (int e = 0; e < n; ++e) {
    gfw[e / L] *= std::fabs(s.fs[e]);
    gbw[e / L] *= std::fabs(s.ff[e] * s.fw[e]);
  }
  double rho = 0;
  for (int p = 0; p < P; ++p) rho = std::max(rho, std::max(gfw[p], gbw[p]));
  if (rho >= 1.0) return P;
  for (int K = 1; K < P; ++K) {
    double worst = 0;
    for (int p0 = 0; p0 + K <= P; ++p0) {
      double pf = 1, pb = 1;
      for (int q = p0; q < p0 + K; ++q) {
        pf *= gfw[q];
        pb *= gbw[q];
      }
      worst = std::max(worst, std::max(pf, pb));
    }
    if (worst < tol * (1.0 - rho)) return K;
  }
  return P;
}

int upload_set(TriSet &s) {
  const size_t n = s.ff.size();
  std::vector<double> h(4 * n);
  for (size_t i = 0; i < n; ++i) {
    h[4 * i + 0] = s.fs[i];
    h[4 * i + 1] = s.fw[i];
    h[4 * i + 2] = s.ff[i] * s.fw[i];
    h[4 * i + 3] = s.sm[i];
  }
  std::vector<float> f(4 * n);
  for (size_t i = 0; i < 4 * n; ++i) f[i] = (float)h[i];
  X3D_HIP(hipMalloc(&s.d64, 4 * n * sizeof(double)));
  X3D_HIP(hipMalloc(&s.d32, 4 * n * sizeof(float)));
  X3D_HIP(hipMemcpy(s.d64, h.data(), 4 * n * sizeof(double), hipMemcpyHostToDevice));
  X3D_HIP(hipMemcpy(s.d32, f.data(), 4 * n * sizeof(float), hipMemcpyHostToDevice));
  // carry-chain depths of the segmented y/z solvers
  s.kd64 = chain_depth(s, SEG_L64, 1e-18);
  s.kd32 = chain_depth(s, SEG_L32, 1e-9);
  return X3D_OK;
}

template <typename T>
Sten<T> cast_sten(const Sten<double> &s) {
  Sten<T> o;
  cast_to(s, o);
  return o;
}
template Sten<double> cast_sten<double>(const Sten<double> &);
template Sten<float> cast_sten<float>(const Sten<double> &);

int build_schemes(Context &c) {
  const x3d_params &p = c.p;
  const int ns[3] = {p.nx, p.ny, p.nz};
  const int b1[3] = {p.nclx1, p.ncly1, p.nclz1};
  const int bn[3] = {p.nclx1 == 0 ? p.nclxn : p.nclxn, p.nclyn, p.nclzn};
  const double lens[3] = {p.xlx, p.yly, p.zlz};
  const char axn[3] = {'x', 'y', 'z'};
  c.sets.clear();
  for (int d = 0; d < 3; ++d) {
    DirCoef &D = c.dir[d];
    D = DirCoef();
    D.n = ns[d];
    D.bc1 = b1[d];
    D.bcn = bn[d];
    if ((D.bc1 == 0) != (D.bcn == 0)) return fail(X3D_ERR_ARG, "periodic BC must be set on both faces");
    if (D.bc1 < 0 || D.bc1 > 2 || D.bcn < 0 || D.bcn > 2) return fail(X3D_ERR_ARG, "ncl must be 0, 1 or 2");
    const bool per = D.bc1 == 0;
    D.nm = per ? D.n : D.n - 1;                        // parameters.f90:272-292
    D.len = lens[d];
    D.d = D.len / (double)D.nm;                        // parameters.f90:294-296
    D.d2 = D.d * D.d;
    if (D.n < 8) return fail(X3D_ERR_UNSUPPORTED, "every direction needs at least 8 points");
    D1 f;
    D2 s;
    ST t;
    X3D_TRY(first_derivative(D.d, D.n, D.bc1, D.bcn, p.ifirstder, f));
    X3D_TRY(second_derivative(D.d2, D.n, D.bc1, D.bcn, p.isecondder, p.nu0nu, p.cnu, s));
    X3D_TRY(interpolation(D.d, D.nm, D.n, D.bc1, D.bcn, p.ifirstder, p.ipinter, t));
    Sten<double> &S = D.s;
    S.afi = f.afi; S.bfi = f.bfi; S.alfai = f.alfai; S.af1 = f.af1; S.bf1 = f.bf1; S.cf1 = f.cf1;
    S.af2 = f.af2; S.afm = f.afm; S.afn = f.afn; S.bfn = f.bfn; S.cfn = f.cfn;
    S.asi = s.asi; S.bsi = s.bsi; S.csi = s.csi; S.dsi = s.dsi; S.alsai = s.alsai;
    S.as1 = s.as1; S.bs1 = s.bs1; S.cs1 = s.cs1; S.ds1 = s.ds1; S.as2 = s.as2; S.as3 = s.as3; S.bs3 = s.bs3;
    S.as4 = s.as4; S.bs4 = s.bs4; S.cs4 = s.cs4; S.astt = s.astt; S.bstt = s.bstt; S.cstt = s.cstt;
    S.ast = s.ast; S.bst = s.bst; S.asm_ = s.asm_; S.asn = s.asn; S.bsn = s.bsn; S.csn = s.csn; S.dsn = s.dsn;
    S.aci = t.aci; S.bci = t.bci; S.alcai = t.alcai; S.aici = t.aici; S.bici = t.bici; S.cici = t.cici;
    S.dici = t.dici; S.ailcai = t.ailcai;
    const std::string a(1, axn[d]);
    // reference variable names (module derivX/Y/Z, module_param.f90:56-75,251-268)
    auto nm = [&](const char *base) {
      std::string b = base;
      return b + a;
    };
    D.arrays = {
        {nm("ff"), f.ff}, {nm("fs"), f.fs}, {nm("fw"), f.fw},
        {nm("ff") + "p", f.ffp}, {nm("fs") + "p", f.fsp}, {nm("fw") + "p", f.fwp},
        {nm("sf"), s.sf}, {nm("ss"), s.ss}, {nm("sw"), s.sw},
        {nm("sf") + "p", s.sfp}, {nm("ss") + "p", s.ssp}, {nm("sw") + "p", s.swp},
        {"cf" + a + "6", t.cfx6}, {"cs" + a + "6", t.csx6}, {"cw" + a + "6", t.cwx6},
        {"cf" + a + "p6", t.cfxp6}, {"cs" + a + "p6", t.csxp6}, {"cw" + a + "p6", t.cwxp6},
        {"cif" + a + "6", t.cifx6}, {"cis" + a + "6", t.cisx6}, {"ciw" + a + "6", t.ciwx6},
        {"cif" + a + "p6", t.cifxp6}, {"cis" + a + "p6", t.cisxp6}, {"ciw" + a + "p6", t.ciwxp6},
        {"cfi6" + (d ? a : std::string()), t.cfi6}, {"csi6" + (d ? a : std::string()), t.csi6},
        {"cwi6" + (d ? a : std::string()), t.cwi6}, {"cfip6" + (d ? a : std::string()), t.cfip6},
        {"csip6" + (d ? a : std::string()), t.csip6}, {"cwip6" + (d ? a : std::string()), t.cwip6},
        {"cifi6" + (d ? a : std::string()), t.cifi6}, {"cisi6" + (d ? a : std::string()), t.cisi6},
        {"ciwi6" + (d ? a : std::string()), t.ciwi6}, {"cifip6" + (d ? a : std::string()), t.cifip6},
        {"cisip6" + (d ? a : std::string()), t.cisip6}, {"ciwip6" + (d ? a : std::string()), t.ciwip6},
    };
    const char *ij = d == 0 ? "i" : (d == 1 ? "j" : "k");
    D.scalars = {
        {std::string("af") + ij + a, f.afi}, {std::string("bf") + ij + a, f.bfi},
        {std::string("alfa") + ij + a, f.alfai}, {"af1" + a, f.af1}, {"bf1" + a, f.bf1}, {"cf1" + a, f.cf1},
        {"af2" + a, f.af2}, {"afm" + a, f.afm}, {"afn" + a, f.afn}, {"bfn" + a, f.bfn}, {"cfn" + a, f.cfn},
        {std::string("as") + ij + a, s.asi}, {std::string("bs") + ij + a, s.bsi},
        {std::string("cs") + ij + a, s.csi}, {std::string("ds") + ij + a, s.dsi},
        {std::string("alsa") + ij + a, s.alsai},
        {"aci" + a + "6", t.aci}, {"bci" + a + "6", t.bci}, {"alcai" + a + "6", t.alcai},
        {"aici" + a + "6", t.aici}, {"bici" + a + "6", t.bici}, {"cici" + a + "6", t.cici},
        {"dici" + a + "6", t.dici}, {"ailcai" + a + "6", t.ailcai},
        {"d" + a, D.d},
    };
    D.set_d1_odd = add_set(c, a + ":ff", f.ff, f.fs, f.fw, per, f.alfai);
    D.set_d1_even = add_set(c, a + ":ffp", f.ffp, f.fsp, f.fwp, per, f.alfai);
    D.set_d2_odd = add_set(c, a + ":sf", s.sf, s.ss, s.sw, per, s.alsai);
    D.set_d2_even = add_set(c, a + ":sfp", s.sfp, s.ssp, s.swp, per, s.alsai);
    D.set_c6 = add_set(c, a + ":cf6", t.cfx6, t.csx6, t.cwx6, per, t.alcai);
    D.set_ci6p = add_set(c, a + ":cifp6", t.cifxp6, t.cisxp6, t.ciwxp6, per, t.ailcai);
    D.set_ci6 = add_set(c, a + ":cif6", t.cifx6, t.cisx6, t.ciwx6, per, t.ailcai);
    if (!per) {
      D.set_cip6 = add_set(c, a + ":cfip6", t.cfip6, t.csip6, t.cwip6, 0, 0.0);
      D.set_ciip6 = add_set(c, a + ":cifip6", t.cifip6, t.cisip6, t.ciwip6, 0, 0.0);
    }
  }
  c.nxm = c.dir[0].nm;
  c.nym = c.dir[1].nm;
  c.nzm = c.dir[2].nm;
  c.xnu = 1.0 / p.re;
  return X3D_OK;
}

int upload_sets(Context &c) {
  for (TriSet &s : c.sets) X3D_TRY(upload_set(s));
  return X3D_OK;
}

void free_sets(Context &c) {
  for (TriSet &s : c.sets) {
    if (s.d64) (void)hipFree(s.d64);
    if (s.d32) (void)hipFree(s.d32);
    s.d64 = nullptr;
    s.d32 = nullptr;
  }
  c.sets.clear();
}

const TriSet *match_set(const Context &c, int dir, int n, const double *ff, const double *fs,
                        const double *fw) {
  const char ax = "xyz"[dir];
  for (const TriSet &s : c.sets) {
    if (s.name[0] != ax || (int)s.ff.size() != n) continue;
    if (std::memcmp(s.ff.data(), ff, n * 8) == 0 && std::memcmp(s.fs.data(), fs, n * 8) == 0 &&
        std::memcmp(s.fw.data(), fw, n * 8) == 0)
      return &s;
  }
  return nullptr;
}

const TriSet *match_set32(const Context &c, int dir, int n, const float *ff, const float *fs,
                          const float *fw) {
  const char ax = "xyz"[dir];
  for (const TriSet &s : c.sets) {
    if (s.name[0] != ax || (int)s.ff.size() != n) continue;
    bool ok = true;
    for (int i = 0; i < n && ok; ++i)
      ok = (float)s.ff[i] == ff[i] && (float)s.fs[i] == fs[i] && (float)s.fw[i] == fw[i];
    if (ok) return &s;
  }
  return nullptr;
}

template <typename T>
int make_line_call(const Context &c, int dir, int op, int npaire, LineCall<T> *out) {
  const DirCoef &D = c.dir[dir];
  const bool per = D.bc1 == 0;
  int set = -1;
  int n_out = D.n;
  switch (op) {
    case OP_DER1: set = npaire == 1 ? D.set_d1_even : D.set_d1_odd; break;
    case OP_DER2: set = npaire == 1 ? D.set_d2_even : D.set_d2_odd; break;
    case OP_DERVP: set = D.set_c6; n_out = D.nm; break;
    case OP_INTERVP: set = D.set_ci6p; n_out = D.nm; break;
    case OP_DERPV: set = per ? D.set_c6 : D.set_cip6; break;
    case OP_INTERPV: set = per ? D.set_ci6 : D.set_ciip6; break;
    default: return fail(X3D_ERR_ARG, "unknown operator");
  }
  if (set < 0) return fail(X3D_ERR_STATE, "coefficient set missing");
  LineCall<T> L;
  L.dir = dir;
  L.op = op;
  L.n_out = n_out;
  if constexpr (sizeof(T) == 8) L.tri = c.sets[set].dev64();
  else L.tri = c.sets[set].dev32();
  L.s = cast_sten<T>(D.s);
  L.bc1 = D.bc1;
  L.bcn = D.bcn;
  L.npaire = npaire;
  *out = L;
  return X3D_OK;
}
template int make_line_call<double>(const Context &, int, int, int, LineCall<double> *);
template int make_line_call<float>(const Context &, int, int, int, LineCall<float> *);

// ---------------------------------------------------------------- stretched y
// stretching.f90:92-318 (alpha != 0: alpha = |xnum/den| > 0 for beta > 0).
// yeta/yetai are the computational coordinates of the velocity and pressure
// points, yp/ypi the physical ones, ppy/ppyi = yly (alpha/pi + sin^2(pi
// eta)/(pi beta)) the metric that the y derivatives multiply by.
int compute_stretching(Context &c) {
  const x3d_params &p = c.p;
  if (p.istret == 0) return X3D_OK;
  if (p.istret < 0 || p.istret > 3) return fail(X3D_ERR_ARG, "istret must be 0, 1, 2 or 3");
  if (!(p.beta > 0.0)) return fail(X3D_ERR_ARG, "stretching needs beta > 0");
  if (p.ncly1 == 0 && p.nclyn == 0)
    return fail(X3D_ERR_UNSUPPORTED, "stretched y needs non-periodic y boundaries");
  const double pi = std::acos(-1.0), yly = p.yly, beta = p.beta;
  const int ny = p.ny, nym = c.dir[1].nm, ist = p.istret;
  const double yinf = -yly / 2.0;
  const double den = 2.0 * beta * yinf;
  const double xnum = -yinf - std::sqrt(pi * pi * beta * beta + yinf * yinf);
  const double alpha = std::fabs(xnum / den);
  c.alpha = alpha;
  auto mapped = [&](double eta) {
    const double den1 = std::sqrt(alpha * beta + 1.0);
    const double xn = den1 / std::sqrt(alpha / pi) / std::sqrt(beta) / std::sqrt(pi);
    const double dn = 2.0 * std::sqrt(alpha / pi) * std::sqrt(beta) * pi * std::sqrt(pi);
    const double den3 = ((std::sin(pi * eta)) * (std::sin(pi * eta)) / beta / pi) + alpha / pi;
    const double den4 = 2.0 * alpha * beta - std::cos(2.0 * pi * eta) + 1.0;
    const double xnum1 = (std::atan(xn * std::tan(pi * eta))) * den4 / den1 / den3 / dn;
    const double cst = std::sqrt(beta) * pi / (2.0 * std::sqrt(alpha) * std::sqrt(alpha * beta + 1.0));
    const double v = eta < 0.5 ? xnum1 - cst : (eta == 0.5 ? 0.0 : xnum1 + cst);
    if (ist == 1) return v - yinf;
    if (ist == 2) return v + yly;
    return (v + yly) * 2.0;
  };
  std::vector<double> yeta(ny), yetai(ny);
  c.yp.assign(ny, 0.0);
  c.ypi.assign(ny, 0.0);
  yeta[0] = ist == 1 ? 0.0 : -0.5;
  for (int j = 1; j < ny; ++j) {
    if (ist == 1) yeta[j] = (double)j * (1.0 / nym);
    else if (ist == 2) yeta[j] = (double)j * (1.0 / nym) - 0.5;
    else yeta[j] = (double)j * (0.5 / nym) - 0.5;
    c.yp[j] = mapped(yeta[j]);
  }
  for (int j = 0; j < ny; ++j) {
    if (ist == 1) yetai[j] = ((double)(j + 1) - 0.5) * (1.0 / nym);
    else if (ist == 2) yetai[j] = ((double)(j + 1) - 0.5) * (1.0 / nym) - 0.5;
    else yetai[j] = ((double)(j + 1) - 0.5) * (0.5 / nym) - 0.5;
    c.ypi[j] = mapped(yetai[j]);
  }
  c.ppy.assign(ny, 0.0);
  c.ppyi.assign(ny, 0.0);
  c.pp2y.assign(ny, 0.0);
  c.pp4y.assign(ny, 0.0);
  for (int j = 0; j < ny; ++j) {
    const double s1 = std::sin(pi * yeta[j]), s2 = std::sin(pi * yetai[j]);
    c.ppy[j] = yly * (alpha / pi + (1.0 / pi / beta) * s1 * s1);
    c.ppyi[j] = yly * (alpha / pi + (1.0 / pi / beta) * s2 * s2);
    // stretching.f90:228-252: pp2y = ppy^2, pp4y = -2/beta cos sin (halved for istret=3)
    c.pp2y[j] = c.ppy[j] * c.ppy[j];
    c.pp4y[j] = (-2.0 / beta * std::cos(pi * yeta[j]) * s1);
    if (ist == 3) c.pp4y[j] = c.pp4y[j] / 2.0;
  }
  return X3D_OK;
}

int build_stretching(Context &c) {
  if (c.p.istret == 0) return X3D_OK;
  X3D_TRY(compute_stretching(c));
  const int ny = c.p.ny;
  std::vector<float> f(ny);
  X3D_HIP(hipMalloc(&c.ppy64, ny * sizeof(double)));
  X3D_HIP(hipMalloc(&c.ppyi64, ny * sizeof(double)));
  X3D_HIP(hipMalloc(&c.ppy32, ny * sizeof(float)));
  X3D_HIP(hipMalloc(&c.ppyi32, ny * sizeof(float)));
  X3D_HIP(hipMemcpy(c.ppy64, c.ppy.data(), ny * sizeof(double), hipMemcpyHostToDevice));
  X3D_HIP(hipMemcpy(c.ppyi64, c.ppyi.data(), ny * sizeof(double), hipMemcpyHostToDevice));
  for (int j = 0; j < ny; ++j) f[j] = (float)c.ppy[j];
  X3D_HIP(hipMemcpy(c.ppy32, f.data(), ny * sizeof(float), hipMemcpyHostToDevice));
  for (int j = 0; j < ny; ++j) f[j] = (float)c.ppyi[j];
  X3D_HIP(hipMemcpy(c.ppyi32, f.data(), ny * sizeof(float), hipMemcpyHostToDevice));
  std::vector<double> pq(ny);
  for (int j = 0; j < ny; ++j) pq[j] = c.pp4y[j] * c.ppy[j];
  X3D_HIP(hipMalloc(&c.pp2y64, ny * sizeof(double)));
  X3D_HIP(hipMalloc(&c.pp4yy64, ny * sizeof(double)));
  X3D_HIP(hipMalloc(&c.pp2y32, ny * sizeof(float)));
  X3D_HIP(hipMalloc(&c.pp4yy32, ny * sizeof(float)));
  X3D_HIP(hipMemcpy(c.pp2y64, c.pp2y.data(), ny * sizeof(double), hipMemcpyHostToDevice));
  X3D_HIP(hipMemcpy(c.pp4yy64, pq.data(), ny * sizeof(double), hipMemcpyHostToDevice));
  for (int j = 0; j < ny; ++j) f[j] = (float)c.pp2y[j];
  X3D_HIP(hipMemcpy(c.pp2y32, f.data(), ny * sizeof(float), hipMemcpyHostToDevice));
  for (int j = 0; j < ny; ++j) f[j] = (float)pq[j];
  X3D_HIP(hipMemcpy(c.pp4yy32, f.data(), ny * sizeof(float), hipMemcpyHostToDevice));
  return X3D_OK;
}

void free_stretching(Context &c) {
  if (c.ppy64) (void)hipFree(c.ppy64);
  if (c.ppyi64) (void)hipFree(c.ppyi64);
  if (c.ppy32) (void)hipFree(c.ppy32);
  if (c.ppyi32) (void)hipFree(c.ppyi32);
  if (c.pp2y64) (void)hipFree(c.pp2y64);
  if (c.pp4yy64) (void)hipFree(c.pp4yy64);
  if (c.pp2y32) (void)hipFree(c.pp2y32);
  if (c.pp4yy32) (void)hipFree(c.pp4yy32);
  c.ppy64 = c.ppyi64 = c.pp2y64 = c.pp4yy64 = nullptr;
  c.ppy32 = c.ppyi32 = c.pp2y32 = c.pp4yy32 = nullptr;
}

}  // namespace x3d
