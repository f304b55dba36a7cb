/*
 * x3d.h -- C-ABI of the MI355X (gfx950) implementation of the Xcompact3d
 * per-timestep hot path: compact finite-difference derivatives, staggered
 * interpolations/derivatives, the spectral pressure Poisson solver and the
 * pencil transposes.
 *
 * Conventions
 *  - Every field argument is a DEVICE pointer to a Fortran-ordered array
 *    (i fastest): element (i,j,k) (1-based) lives at (i-1)+n1*((j-1)+n2*(k-1)).
 *    Callers whose arrays live in host memory use the x3d_host_* entries.
 *  - Scalar arguments of the Fortran-mirror entries are passed by reference,
 *    as a Fortran caller passes them (bind(C) without VALUE).
 *  - The coefficient arrays (ff,fs,fw ...) are HOST pointers, exactly the
 *    module arrays the reference callers pass; the library matches them to
 *    the device copies it built in x3d_init (or uploads them if they are not
 *    one of those sets).
 *  - The scratch arguments r and s of the reference (rx, sx) are accepted and
 *    ignored: the result lives only in t, as in the reference.
 *  - Every entry returns 0 on success and a non-zero x3d_status otherwise;
 *    x3d_last_error() describes the last failure of the calling thread.
 *  - All work is enqueued on the library stream (x3d_set_stream); the
 *    Fortran-mirror entries synchronise that stream before returning so that
 *    their semantics equal the reference subroutine call; the *_async entries
 *    do not.
 *  - The library is not re-entrant: one context per process, as the
 *    reference's module state (module_param.f90, variables.f90).
 */
#ifndef X3D_H
#define X3D_H

#ifdef __cplusplus
extern "C" {
#endif

enum x3d_status {
  X3D_OK = 0,
  X3D_ERR_ARG = 1,        /* bad argument (size, bc, pointer) */
  X3D_ERR_STATE = 2,      /* x3d_init not called, or called twice */
  X3D_ERR_HIP = 3,        /* a HIP runtime call failed */
  X3D_ERR_UNSUPPORTED = 4, /* option the library does not implement */
  X3D_ERR_IO = 5           /* a checkpoint file could not be written or read */
};

/* Subset of the reference's namelist parameters that the hot path reads
 * (parameters.f90:44-86; derived sizes parameters.f90:272-304). */
typedef struct x3d_params {
  int nx, ny, nz;                  /* velocity-grid points */
  double xlx, yly, zlz;            /* domain lengths */
  int nclx1, nclxn, ncly1, nclyn, nclz1, nclzn; /* 0 periodic, 1 free-slip, 2 Dirichlet */
  int ifirstder;                   /* 4 = 6th-order compact (1 = 2nd-order central) */
  int isecondder;                  /* 4 = 6th-order compact, 5 = hyperviscous */
  int ipinter;                     /* 1,2,3: mid-point interpolation family (default 3) */
  double re;                       /* xnu = 1/re */
  double dt;
  int itimescheme;                 /* 3 = AB3 (the only scheme of the step driver) */
  int istret;                      /* 0 uniform; 1, 2, 3 stretched y (stretching.f90) */
  double nu0nu, cnu;               /* isecondder=5 parameters */
  double beta;                     /* stretching parameter (istret != 0) */
} x3d_params;

/* ---- context -------------------------------------------------------------
 * x3d_init replaces schemes() (schemes.f90:7-409) and
 * decomp_2d_poisson_init() (poisson.f90:73-255): it computes every
 * coefficient set in FP64 on the host, uploads FP64 and FP32 copies, and
 * builds the modified-wavenumber table of waves() (poisson.f90:1530-1810). */
int x3d_init(const x3d_params *p);
int x3d_finalize(void);
const char *x3d_last_error(void);
int x3d_set_stream(void *hip_stream);      /* NULL = default stream */
int x3d_synchronize(void);
int x3d_device_name(char *buf, int len);

/* Copy a coefficient array computed by x3d_init into host memory, so that a
 * caller can hand the same arrays back (the reference's module arrays).
 * name is the reference variable name: "ffx","fsx","fwx","ffxp",...,
 * "sfx",...,"cfx6","csx6","cwx6","cifxp6",... ("y"/"z" variants likewise),
 * and with istret != 0 the stretching.f90 arrays "yp","ypi","ppy","ppyi",
 * "pp2y","pp4y" (stretching.f90:92-318).  Returns the array length in *n. */
int x3d_get_coefficients(const char *name, double *out, int cap, int *n);
/* Scalar coefficient by reference name: "afix","bfix","alfaix","asix",...,
 * "acix6","alcaix6","aicix6","ailcaix6", "xnu", "dx", ... */
int x3d_get_scalar(const char *name, double *out);
/* Host-only evaluation of schemes() for the parameters p (no device, no
 * context): an array or scalar by the same names as above. */
int x3d_schemes_host(const x3d_params *p, const char *name, double *out, int cap, int *n);

/* ---- compact derivatives -------------------------------------------------
 * Replaces the procedure pointers derx/dery/derz/derxx/deryy/derzz bound in
 * schemes.f90:31-66 with the abstract interfaces DERIVATIVE_X/Y/YY/Z of
 * module_param.f90:135-168; one entry per BC variant of derive.f90:
 *   derx_00 derive.f90:7    derx_11 :68    derx_12 :141   derx_21 :211   derx_22 :281
 *   dery_00 :325            dery_11 :423   dery_12 :545   dery_21 :664   dery_22 :783
 *   derz_00 :858            derz_11 :951   derz_12 :1064  derz_21 :1176  derz_22 :1288
 *   derxx_00 :1354 derxx_11 :1481 derxx_12 :1666 derxx_21 :1822 derxx_22 :1978
 *   deryy_00 :2052 deryy_11 :2207 deryy_12 :2433 deryy_21 :2631 deryy_22 :2829
 *   derzz_00 :2923 derzz_11 :3082 derzz_12 :3307 derzz_21 :3503 derzz_22 :3699
 * t,u device; r,s ignored; ff,fs,fw host; pp (y, first derivative only) host. */
#define X3D_DECL_DER_X(name, T)                                                  \
  int name(T *t, const T *u, T *r, T *s, const T *ff, const T *fs, const T *fw,  \
           const int *nx, const int *ny, const int *nz, const int *npaire,       \
           const T *lind);
#define X3D_DECL_DER_Y(name, T)                                                  \
  int name(T *t, const T *u, T *r, T *s, const T *ff, const T *fs, const T *fw,  \
           const T *pp, const int *nx, const int *ny, const int *nz,             \
           const int *npaire, const T *lind);

#define X3D_DECL_DER_FAMILY(T, SFX)                                              \
  X3D_DECL_DER_X(x3d_derx_00##SFX, T) X3D_DECL_DER_X(x3d_derx_11##SFX, T)        \
  X3D_DECL_DER_X(x3d_derx_12##SFX, T) X3D_DECL_DER_X(x3d_derx_21##SFX, T)        \
  X3D_DECL_DER_X(x3d_derx_22##SFX, T)                                            \
  X3D_DECL_DER_Y(x3d_dery_00##SFX, T) X3D_DECL_DER_Y(x3d_dery_11##SFX, T)        \
  X3D_DECL_DER_Y(x3d_dery_12##SFX, T) X3D_DECL_DER_Y(x3d_dery_21##SFX, T)        \
  X3D_DECL_DER_Y(x3d_dery_22##SFX, T)                                            \
  X3D_DECL_DER_X(x3d_derz_00##SFX, T) X3D_DECL_DER_X(x3d_derz_11##SFX, T)        \
  X3D_DECL_DER_X(x3d_derz_12##SFX, T) X3D_DECL_DER_X(x3d_derz_21##SFX, T)        \
  X3D_DECL_DER_X(x3d_derz_22##SFX, T)                                            \
  X3D_DECL_DER_X(x3d_derxx_00##SFX, T) X3D_DECL_DER_X(x3d_derxx_11##SFX, T)      \
  X3D_DECL_DER_X(x3d_derxx_12##SFX, T) X3D_DECL_DER_X(x3d_derxx_21##SFX, T)      \
  X3D_DECL_DER_X(x3d_derxx_22##SFX, T)                                           \
  X3D_DECL_DER_X(x3d_deryy_00##SFX, T) X3D_DECL_DER_X(x3d_deryy_11##SFX, T)      \
  X3D_DECL_DER_X(x3d_deryy_12##SFX, T) X3D_DECL_DER_X(x3d_deryy_21##SFX, T)      \
  X3D_DECL_DER_X(x3d_deryy_22##SFX, T)                                           \
  X3D_DECL_DER_X(x3d_derzz_00##SFX, T) X3D_DECL_DER_X(x3d_derzz_11##SFX, T)      \
  X3D_DECL_DER_X(x3d_derzz_12##SFX, T) X3D_DECL_DER_X(x3d_derzz_21##SFX, T)      \
  X3D_DECL_DER_X(x3d_derzz_22##SFX, T)

X3D_DECL_DER_FAMILY(double, )
X3D_DECL_DER_FAMILY(float, _f32)

#define X3D_DECL_DER_FAMILY_HOST(T, SFX)                                         \
  X3D_DECL_DER_X(x3d_host_derx_00##SFX, T) X3D_DECL_DER_X(x3d_host_derx_11##SFX, T)   \
  X3D_DECL_DER_X(x3d_host_derx_12##SFX, T) X3D_DECL_DER_X(x3d_host_derx_21##SFX, T)   \
  X3D_DECL_DER_X(x3d_host_derx_22##SFX, T)                                       \
  X3D_DECL_DER_Y(x3d_host_dery_00##SFX, T) X3D_DECL_DER_Y(x3d_host_dery_11##SFX, T)   \
  X3D_DECL_DER_Y(x3d_host_dery_12##SFX, T) X3D_DECL_DER_Y(x3d_host_dery_21##SFX, T)   \
  X3D_DECL_DER_Y(x3d_host_dery_22##SFX, T)                                       \
  X3D_DECL_DER_X(x3d_host_derz_00##SFX, T) X3D_DECL_DER_X(x3d_host_derz_11##SFX, T)   \
  X3D_DECL_DER_X(x3d_host_derz_12##SFX, T) X3D_DECL_DER_X(x3d_host_derz_21##SFX, T)   \
  X3D_DECL_DER_X(x3d_host_derz_22##SFX, T)                                       \
  X3D_DECL_DER_X(x3d_host_derxx_00##SFX, T) X3D_DECL_DER_X(x3d_host_derxx_11##SFX, T) \
  X3D_DECL_DER_X(x3d_host_derxx_12##SFX, T) X3D_DECL_DER_X(x3d_host_derxx_21##SFX, T) \
  X3D_DECL_DER_X(x3d_host_derxx_22##SFX, T)                                      \
  X3D_DECL_DER_X(x3d_host_deryy_00##SFX, T) X3D_DECL_DER_X(x3d_host_deryy_11##SFX, T) \
  X3D_DECL_DER_X(x3d_host_deryy_12##SFX, T) X3D_DECL_DER_X(x3d_host_deryy_21##SFX, T) \
  X3D_DECL_DER_X(x3d_host_deryy_22##SFX, T)                                      \
  X3D_DECL_DER_X(x3d_host_derzz_00##SFX, T) X3D_DECL_DER_X(x3d_host_derzz_11##SFX, T) \
  X3D_DECL_DER_X(x3d_host_derzz_12##SFX, T) X3D_DECL_DER_X(x3d_host_derzz_21##SFX, T) \
  X3D_DECL_DER_X(x3d_host_derzz_22##SFX, T)

/* ---- staggered operators (velocity grid n <-> pressure grid nm) ----------
 * Same argument order as the reference external subroutines:
 *   derxvp   derive.f90:3796  (tx,ux,rx,sx,cfx6,csx6,cwx6,nx,nxm,ny,nz,npaire)
 *   interxvp derive.f90:3911  (tx,ux,rx,sx,cifx6,cisx6,ciwx6,nx,nxm,ny,nz,npaire)
 *   derxpv   derive.f90:4041  (tx,ux,rx,sx,cfi6,csi6,cwi6,cfx6,csx6,cwx6,nxm,nx,ny,nz,npaire)
 *   interxpv derive.f90:4126  (tx,ux,rx,sx,cifi6,cisi6,ciwi6,cifx6,cisx6,ciwx6,nxm,nx,ny,nz,npaire)
 *   interyvp :4265 deryvp :4442 (+ppyi) interypv :4587 derypv :4775 (+ppy)
 *   derzvp :4920 interzvp :5105 derzpv :5287 interzpv :5426
 */
#define X3D_DECL_STAG_P(P, T, SFX)                                                    \
  int P##derxvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs,    \
                      const T *cw, const int *nx, const int *nxm, const int *ny, \
                      const int *nz, const int *npaire);                         \
  int P##interxvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs,  \
                        const T *cw, const int *nx, const int *nxm,              \
                        const int *ny, const int *nz, const int *npaire);        \
  int P##derxpv##SFX(T *t, const T *u, T *r, T *s, const T *cf1, const T *cs1,  \
                      const T *cw1, const T *cf2, const T *cs2, const T *cw2,    \
                      const int *nxm, const int *nx, const int *ny,              \
                      const int *nz, const int *npaire);                         \
  int P##interxpv##SFX(T *t, const T *u, T *r, T *s, const T *cf1,              \
                        const T *cs1, const T *cw1, const T *cf2, const T *cs2,  \
                        const T *cw2, const int *nxm, const int *nx,             \
                        const int *ny, const int *nz, const int *npaire);        \
  int P##interyvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs,  \
                        const T *cw, const int *nx, const int *ny,               \
                        const int *nym, const int *nz, const int *npaire);       \
  int P##deryvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs,    \
                      const T *cw, const T *ppyi, const int *nx, const int *ny,  \
                      const int *nym, const int *nz, const int *npaire);         \
  int P##interypv##SFX(T *t, const T *u, T *r, T *s, const T *cf1,              \
                        const T *cs1, const T *cw1, const T *cf2, const T *cs2,  \
                        const T *cw2, const int *nx, const int *nym,             \
                        const int *ny, const int *nz, const int *npaire);        \
  int P##derypv##SFX(T *t, const T *u, T *r, T *s, const T *cf1, const T *cs1,  \
                      const T *cw1, const T *cf2, const T *cs2, const T *cw2,    \
                      const T *ppy, const int *nx, const int *nym,               \
                      const int *ny, const int *nz, const int *npaire);          \
  int P##derzvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs,    \
                      const T *cw, const int *nx, const int *ny, const int *nz,  \
                      const int *nzm, const int *npaire);                        \
  int P##interzvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs,  \
                        const T *cw, const int *nx, const int *ny,               \
                        const int *nz, const int *nzm, const int *npaire);       \
  int P##derzpv##SFX(T *t, const T *u, T *r, T *s, const T *cf1, const T *cs1,  \
                      const T *cw1, const T *cf2, const T *cs2, const T *cw2,    \
                      const int *nx, const int *ny, const int *nzm,              \
                      const int *nz, const int *npaire);                         \
  int P##interzpv##SFX(T *t, const T *u, T *r, T *s, const T *cf1,              \
                        const T *cs1, const T *cw1, const T *cf2, const T *cs2,  \
                        const T *cw2, const int *nx, const int *ny,              \
                        const int *nzm, const int *nz, const int *npaire);
#define X3D_DECL_STAG(T, SFX) X3D_DECL_STAG_P(x3d_, T, SFX)
#define X3D_DECL_STAG_HOST(T, SFX) X3D_DECL_STAG_P(x3d_host_, T, SFX)

X3D_DECL_STAG(double, )
X3D_DECL_STAG(float, _f32)

/* ---- generic asynchronous line operator ----------------------------------
 * The operator behind every entry above, with the coefficient set the
 * reference callers pass (ff*p for npaire=1, ...), enqueued on the library
 * stream without synchronising.  dir 0/1/2 = x/y/z; op 0 der1, 1 der2,
 * 2 dervp, 3 intervp, 4 derpv, 5 interpv; precision 0 = FP64, 1 = FP32;
 * (n1,n2,n3) are the extents of the input array.  On a stretched y mesh
 * (istret != 0) the y operators multiply by the metric of x3d_init as the
 * named entries do: der1 and derpv by ppy, dervp by ppyi. */
int x3d_line_op_async(int dir, int op, int npaire, int precision, const void *u, void *t, int n1, int n2,
                      int n3);

/* ---- pressure Poisson -----------------------------------------------------
 * Replaces the procedure pointer `poisson` (poisson.f90:57-65) bound by
 * decomp_2d_poisson_init to poisson_000 (poisson.f90:298), poisson_100
 * (poisson.f90:413), poisson_010 (poisson.f90:665; with istret != 0 its
 * matrice_refinement branch, poisson.f90:846-895) or poisson_11x
 * (poisson.f90:1019).  rhs is the
 * device pressure-grid field (nxm,nym,nzm), overwritten with the solution,
 * exactly like poisson_xxx(rhs).  The _f32 entry keeps FP32 storage and FFT
 * arithmetic and divides by the modified wavenumber (or applies the
 * pentadiagonal factors) in FP64. */
int x3d_poisson(double *rhs, const int *n1, const int *n2, const int *n3);
int x3d_poisson_f32(float *rhs, const int *n1, const int *n2, const int *n3);
/* Enqueue-only variants for whole-step callers (no counterpart in the
 * reference, whose calls are synchronous): the work is queued on the library
 * stream and the call returns at once.  precision 0 = FP64, 1 = FP32.
 * x3d_transpose_async stands for all four transposes (one GPU: a copy). */
int x3d_poisson_async(int precision, void *rhs, int n1, int n2, int n3);
/* x3d_poisson_async without allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The same
 * kernels as x3d_poisson_async, with the spectral work array taken from a
 * file-scope device array of 160 MiB (167772160 bytes) instead of a buffer
 * the library allocates: it covers (nxm/2+1)*nym*nzm complex words of the
 * precision, e.g. 256^3 in FP64 or 256x512x512 in FP32; larger grids return
 * X3D_ERR_UNSUPPORTED.  The call allocates nothing (x3d_init must have run:
 * it records the array's device address).  Every call on one GPU shares
 * that GPU's array, so calls that overlap on one GPU (other streams, other
 * host threads) race. */
int x3d_poisson_async_noalloc(int precision, void *rhs, int n1, int n2, int n3);
int x3d_transpose_async(int precision, const void *src, void *dst, int n1, int n2, int n3);

/* ---- 3-D FFT ---------------------------------------------------------------
 * decomp_2d_fft_3d of 2DECOMP&FFT v2.0.4 (external; called at
 * poisson.f90:330, 405, 455, 638, 705, 994) after
 * decomp_2d_fft_init(PHYSICAL_IN_Z): the unnormalised transform of the real
 * pressure-grid field in (n1,n2,n3) = (nxm,nym,nzm) to its complex spectrum
 * out (n1, n2, n3/2+1), i fastest, interleaved (re, im) as Fortran complex,
 * with exp(-i k.x) (r2c) and back with exp(+i k.x) (c2r; c2r(r2c(f)) =
 * n1 n2 n3 f).  The c2r input must be the spectrum of a real field (as the
 * reference's is); it is not modified.  Device arrays, out of place,
 * blocking.  Hand-written radix kernels, power-of-two extents as for the
 * Poisson solver. */
int x3d_fft3d_r2c(const double *in, void *out, const int *n1, const int *n2, const int *n3);
int x3d_fft3d_c2r(const void *in, double *out, const int *n1, const int *n2, const int *n3);
int x3d_fft3d_r2c_f32(const float *in, void *out, const int *n1, const int *n2, const int *n3);
int x3d_fft3d_c2r_f32(const void *in, float *out, const int *n1, const int *n2, const int *n3);

/* ---- pencil transposes ----------------------------------------------------
 * 2DECOMP&FFT v2.0.4 transpose_x_to_y / y_to_z / z_to_y / y_to_x (external;
 * call sites e.g. transeq.f90:163-165, navier.f90:316-318).  On one GPU every
 * pencil is the whole (n1,n2,n3) Fortran-ordered array, so each transpose is
 * a device-resident copy of n1*n2*n3 words. */
int x3d_transpose_x_to_y(const double *src, double *dst, const int *n1, const int *n2, const int *n3);
int x3d_transpose_y_to_z(const double *src, double *dst, const int *n1, const int *n2, const int *n3);
int x3d_transpose_z_to_y(const double *src, double *dst, const int *n1, const int *n2, const int *n3);
int x3d_transpose_y_to_x(const double *src, double *dst, const int *n1, const int *n2, const int *n3);
/* Single-precision builds of the reference (mytype = real(4)). */
int x3d_transpose_x_to_y_f32(const float *src, float *dst, const int *n1, const int *n2, const int *n3);
int x3d_transpose_y_to_z_f32(const float *src, float *dst, const int *n1, const int *n2, const int *n3);
int x3d_transpose_z_to_y_f32(const float *src, float *dst, const int *n1, const int *n2, const int *n3);
int x3d_transpose_y_to_x_f32(const float *src, float *dst, const int *n1, const int *n2, const int *n3);
/* LDS-tiled layout change dst(j,i,k) = src(i,j,k) of an (n1,n2,n3) FP64
 * array, out of place: the reorder a y-pencil needs when it is stored with
 * y fastest.  No counterpart in the reference (its pencils keep i fastest). */
int x3d_reorder_ji(const double *src, double *dst, const int *n1, const int *n2, const int *n3);

/* ---- host-memory entries -------------------------------------------------
 * x3d_host_<name> takes exactly the arguments of x3d_<name> above, with the
 * field arrays (t, u; rhs; in, out; src, dst) in host memory, as the
 * reference's allocatables are (variables.f90:127-187): the input is copied
 * to HBM, the device entry runs, the output is copied back.  For Fortran
 * callers that keep their fields on the host; the PCIe copies make them the
 * slow path (x3d_solver_* keeps everything resident).  The transposes are
 * identity copies in host memory. */
X3D_DECL_DER_FAMILY_HOST(double, )
X3D_DECL_DER_FAMILY_HOST(float, _f32)
X3D_DECL_STAG_HOST(double, )
X3D_DECL_STAG_HOST(float, _f32)
int x3d_host_poisson(double *rhs, const int *n1, const int *n2, const int *n3);
int x3d_host_poisson_f32(float *rhs, const int *n1, const int *n2, const int *n3);
int x3d_host_fft3d_r2c(const double *in, void *out, const int *n1, const int *n2, const int *n3);
int x3d_host_fft3d_c2r(const void *in, double *out, const int *n1, const int *n2, const int *n3);
int x3d_host_fft3d_r2c_f32(const float *in, void *out, const int *n1, const int *n2, const int *n3);
int x3d_host_fft3d_c2r_f32(const void *in, float *out, const int *n1, const int *n2, const int *n3);
int x3d_host_transpose_x_to_y(const double *src, double *dst, const int *n1, const int *n2, const int *n3);
int x3d_host_transpose_y_to_z(const double *src, double *dst, const int *n1, const int *n2, const int *n3);
int x3d_host_transpose_z_to_y(const double *src, double *dst, const int *n1, const int *n2, const int *n3);
int x3d_host_transpose_y_to_x(const double *src, double *dst, const int *n1, const int *n2, const int *n3);
int x3d_host_transpose_x_to_y_f32(const float *src, float *dst, const int *n1, const int *n2, const int *n3);
int x3d_host_transpose_y_to_z_f32(const float *src, float *dst, const int *n1, const int *n2, const int *n3);
int x3d_host_transpose_z_to_y_f32(const float *src, float *dst, const int *n1, const int *n2, const int *n3);
int x3d_host_transpose_y_to_x_f32(const float *src, float *dst, const int *n1, const int *n2, const int *n3);

/* ---- device-resident time step (navier.f90 / transeq.f90 / time_integrators.f90)
 * A solver owns the velocity (ux,uy,uz), the AB3 history (dux1..3 of
 * time_integrators.f90:85-100) and all work arrays in HBM.
 * precision: 0 = FP64 everywhere; 1 = FP32 bulk (FP32 derivatives, FFTs,
 * convective products) with FP64 velocity accumulation, FP64 Poisson divide
 * and FP64 diagnostics reductions.
 * For nx in {64,128,256,512} and ny, nz multiples of 16 up to 512 the
 * momentum right-hand side and AB3 run as nine fused sweeps and the
 * divergence / gradient + cor_vel as six; other grids (or the environment
 * variable X3D_UNFUSED_MOMENTUM set at create time) use one launch per
 * derivative.  Both paths compute the same quantities.
 * Boundaries: periodic or free-slip (ncl=1 at both ends) per direction, and
 * in y also no-slip walls (ncl=2, either or both faces), with the Poisson
 * solver of the grid (poisson_000, poisson_010 with its stretched branch,
 * poisson_11x).  pre_correc sets the zero normal velocity on free-slip faces
 * (navier.f90:600-770) and the wall velocity on no-slip faces from the
 * pressure gradient gradp stored there (navier.f90:456-479, 616-690).
 * A stretched y mesh (istret != 0) multiplies the y derivatives by ppy /
 * ppyi and forms the y viscous term as pp2y deryy - pp4y dery
 * (transeq.f90:337-370).
 * Time schemes: AB3 (itimescheme=3) and the low-storage RK3 (itimescheme=5,
 * time_integrators.f90:151-157, three projected sub-steps per step); the
 * fused sweeps run AB3 on periodic boxes. */
typedef struct x3d_solver x3d_solver;
int x3d_solver_create(int precision, x3d_solver **out);
int x3d_solver_destroy(x3d_solver *s);
/* Taylor-Green initial condition of Case-TGV.f90:64-96 (iin=1). */
int x3d_solver_init_tgv(x3d_solver *s);
/* Set the velocity from device FP64 arrays (nx,ny,nz) and reset itime. */
int x3d_solver_set_velocity(x3d_solver *s, const double *ux, const double *uy, const double *uz);
/* Copy the FP64 velocity to device arrays. */
int x3d_solver_get_velocity(x3d_solver *s, double *ux, double *uy, double *uz);
/* nsteps iterations of xcompact3d.f90:29-102 (transeq RHS, AB3, divergence,
 * Poisson, gradp, cor_vel). */
int x3d_solver_step(x3d_solver *s, int nsteps);
/* (eek, eps, eps2, enst) of postprocess_tgv (Case-TGV.f90:246-377), FP64. */
int x3d_solver_tgv_stats(x3d_solver *s, double out4[4]);
/* Max and mean |div u| on the pressure grid after the projection
 * (navier.f90:346-361 with nlock=2 semantics, FP64). */
int x3d_solver_divergence_stats(x3d_solver *s, double out2[2]);
/* The statistics the reference's divergence computes on every call
 * (navier.f90:348-361, nlock=1): max and mean |.| of div u* on the pressure
 * grid, u* the velocity before the pressure correction of the last
 * (sub-)step, in the compute precision of the step (tmax starting at -1609
 * as there).  The step reduces them per block as it goes (in the z
 * divergence sweep on the fused path); this call finishes the reduction.
 * X3D_ERR_STATE before the first step. */
int x3d_solver_divu_stats(x3d_solver *s, double out2[2]);
int x3d_solver_itime(x3d_solver *s, int *itime);
/* Channel initial condition of init_channel (Case-Channel.f90:96-128, iin=1,
 * idir_stream=1): u = init_noise um (2r-1) + 1 - y^2, v, w = init_noise um
 * (2r-1), um = exp(-0.2 y^2), y the wall-normal coordinate (yp on a
 * stretched mesh) minus yly/2.  The reference draws r with the compiler's
 * random_number; here r is a counter-based SplitMix64 number of (seed, the
 * component, the point index i + nx (j + ny k)), identical on every device
 * and in the oracle.  seed < 2^24.  Resets itime and the AB3 history. */
int x3d_solver_init_channel(x3d_solver *s, double init_noise, unsigned long long seed);
/* Streamwise forcing of the channel, applied at every (sub-)step:
 * mode 0 none; 1 constant flow rate, channel_cfr(ux, value) before the
 * right-hand side (Case-Channel.f90:165-171, 220-261; value 2/3 in the
 * reference); 2 constant pressure gradient, value added to the x right-hand
 * side (cpg, fcpg of Case-Channel.f90:403-410). */
int x3d_solver_set_forcing(x3d_solver *s, int mode, double value);
/* Spin-up rotation of the channel (momentum_forcing_channel,
 * Case-Channel.f90:413-418; wrotation and spinup_time of the &Statistics
 * namelist): while the step index (1 at the first step, as itime with
 * ifirst = 1) is below spinup_time, every (sub-)step adds -wrotation * uy
 * to the x and +wrotation * ux to the y right-hand side.  spinup_time 0
 * turns it off.  Per-operator step only (bounded y, or RK3). */
int x3d_solver_set_rotation(x3d_solver *s, double wrotation, int spinup_time);
/* Bulk velocity ub = dy / (yly nx nz) sum ux / ppy of channel_cfr
 * (Case-Channel.f90:234-246), FP64. */
int x3d_solver_bulk_velocity(x3d_solver *s, double *ub);
/* restart(..., iresflg=1) of tools.f90:205-383: writes the checkpoint of
 * the solver's state after its last step into <rundir>/checkpoint/, one file
 * per variable with the reference's names (ux, uy, uz; dux-2, duy-2, duz-2,
 * dux-3, duy-3, duz-3 for AB3; pp on the pressure grid), each the global
 * array in Fortran order as native FP64 words (decomp_2d_write_one,
 * reduce_prec=.false.), and <rundir>/restart.info with the &Time (tfield =
 * t, itime) and &NumParam namelists in the reference's formats
 * (tools.f90:343-373).  X3D_ERR_STATE before the first step. */
int x3d_solver_checkpoint(x3d_solver *s, const char *rundir, double t);
/* restart(..., iresflg=0) of tools.f90:384-500: reads what
 * x3d_solver_checkpoint wrote into a solver of the same grid, scheme and
 * precision: velocity, AB3 history, pressure, itime from the &Time namelist
 * (its tfield returned in *t), and with walls in y the wall planes of the
 * pressure gradient, reconstructed from pp as gradp does (tools.f90:493-500).
 * The next step continues the run bit for bit.  X3D_ERR_IO for a missing
 * file or one whose size does not match the grid. */
int x3d_solver_restart(x3d_solver *s, const char *rundir, double *t);
/* Instrumentation, no counterpart in the reference: enqueue nrep times the
 * six fused y/z momentum sweeps of one step (z then y for each component,
 * on the current velocity, into the solver's work buffer; the velocity and
 * the AB3 history are not touched) on the library stream, without
 * synchronizing.  *bytes_per_launch receives the HBM bytes one launch moves
 * on average (u_d when it differs from u_c, u_c, the accumulator when
 * present, the output).  Fused solvers only. */
int x3d_solver_momentum_yz(x3d_solver *s, int nrep, double *bytes_per_launch);

/* ---- LES subgrid stress (les_models.f90) ----------------------------------
 * model 0: none (DNS); 1: Smagorinsky, jles=1 of compute_SGS
 * (les_models.f90:96-165) with smagcst = cs and iconserv 0
 * (sgs_mom_nonconservative, :1091-1250) or 1 (sgs_mom_conservative,
 * :1327-1456); 2: the learned closure network `mlp` (no counterpart in the
 * reference) -- T_ij from the nine resolved gradients over gscale and the
 * filter width over delta0, divergence with the conservative parities, and
 * the Smagorinsky stress of constant cs wherever T_ij S_ij < -beta 2 nut S:S;
 * 3: the same network, keeping only the dissipative part of its stress
 * along the strain, T = 2 nu_t S with nu_t = max(T:S, 0) / (2 S:S).
 * The force is added to the momentum right-hand side before AB3, as
 * transeq.f90:498-503 adds sgsx1..sgsz1.  The solver does not own mlp. */
typedef struct x3d_mlp x3d_mlp;
int x3d_solver_set_les(x3d_solver *s, int model, double cs, int iconserv, double beta, x3d_mlp *mlp,
                       double gscale, double delta0);
/* SGS force of the current velocity into device FP64 arrays (nx,ny,nz). */
int x3d_solver_sgs(x3d_solver *s, double *sgsx, double *sgsy, double *sgsz);
/* Points where the learned closure fell back to Smagorinsky since the last
 * call (the counter is reset). */
int x3d_solver_les_fallbacks(x3d_solver *s, long long *count);

/* ---- closure network (BF16 MFMA, FP32 accumulation) ------------------------
 * 10 -> 64 -> 64 -> 64 -> 6 with GELU; x3d_mlp_nparams() FP32 parameters in
 * the order W1 (64x10, row-major), b1, W2 (64x64), b2, W3, b3, W4 (6x64), b4.
 * Feature rows (n x 10, device FP32): g_xx g_yx g_zx g_xy g_yy g_zy g_xz g_yz
 * g_zz (g_ij = d u_i / d x_j) over gscale, then delta/delta0.  Output rows
 * (n x 6): T_xx T_yy T_zz T_xy T_xz T_yz over delta^2 gscale^2.
 * Training: mean squared error over the n x 6 entries, Adam (lr, beta1 0.9,
 * beta2 0.999, eps 1e-8) on FP32 master weights.  Not re-entrant. */
int x3d_mlp_nparams(void);
int x3d_mlp_create(x3d_mlp **out);
int x3d_mlp_destroy(x3d_mlp *m);
int x3d_mlp_set_params(x3d_mlp *m, const float *host_params);   /* also resets Adam */
int x3d_mlp_get_params(x3d_mlp *m, float *host_params);
int x3d_mlp_forward(x3d_mlp *m, const float *x, float *y, int n);
/* One Adam step on a batch; *loss (host, may be NULL) = the batch MSE. */
int x3d_mlp_train_step(x3d_mlp *m, const float *x, const float *t, int n, float lr, float *loss);
/* Gradient of the batch MSE into host_grad (x3d_mlp_nparams() floats), no update. */
int x3d_mlp_gradient(x3d_mlp *m, const float *x, const float *t, int n, float *host_grad, float *loss);
/* T_ij (device, precision 0 FP64 / 1 FP32) from the nine gradient fields g
 * (order as the features), with the Smagorinsky fallback; project 0 keeps
 * the network's stress (the solver's LES model 2), 1 its eddy-viscosity
 * projection T = 2 nu_t S, nu_t = max(T:S, 0) / (2 S:S) (model 3);
 * *nfallback (host, may be NULL) counts the fallback points. */
int x3d_mlp_closure(x3d_mlp *m, int precision, const void *const *g, void *const *t, long long n, double gscale,
                    double delta, double delta0, double cs, double beta, int project, long long *nfallback);

/* Top-hat filter of a device FP64 velocity (nx,ny,nz) over f^3 cells onto
 * (nx/f, ny/f, nz/f): ub[3] the filtered velocity, tau[6] (xx yy zz xy xz yz)
 * = bar(u_i u_j) - bar(u_i) bar(u_j).  Training data for the closure; no
 * counterpart in the reference. */
int x3d_box_filter(const double *ux, const double *uy, const double *uz, int nx, int ny, int nz, int f,
                   double *const *ub, double *const *tau);

#ifdef __cplusplus
}
#endif
#endif /* X3D_H */
