/*
 * tck_step.hip -- FP64 stencil kernels of one RK3 substep (gfx950).
 *
 *  tck_predict      fluid_predict_field (src/fluid/predict/main.c:35-110) for ux, uy, uz
 *                   and T in ONE sweep: the 11 velocity-gradient arrays of the reference
 *                   (predict/l*.c) are formed on the fly from the velocity neighbours, every
 *                   explicit (srca) and implicit (srcg) term is accumulated in the order of
 *                   compute_rhs_* (ux.c:402-439, uy.c:427-465, uz.c:326-357, t.c:247-272),
 *                   and the RK increment (ux.c:476-494) is formed in registers.  It writes
 *                   the explicit RHS for the next substep and either the increment (when a
 *                   line solve follows) or the new field with its periodic halos.
 *  tck_correct      correct/*.c and update_pressure.c fused: u -= g dt grad psi, p += psi
 *                   (- implicit-diffusion terms), with halos and the Neumann wall copy of p;
 *                   optionally the CFL minimum of decide_dt.c:32-80 on the corrected field.
 *  tck_dt_*         decide_dt (decide_dt.c:32-160) as a device min reduction + finalise.
 *
 * Every arithmetic expression keeps the reference's operand order; built with
 * -ffp-contract=off the stencil terms round exactly as the reference's do.
 */
#include "tck_common.h"

/* Wray RK3, src/runge_kutta.c:8-15 */
__constant__ double c_rka[3] = {32. / 60., 25. / 60., 45. / 60.};
__constant__ double c_rkb[3] = { 0. / 60., -17. / 60., -25. / 60.};

struct PredictArgs {
  tc_geom_t g;
  tc_coef_t c;
  tc_flags_t f;
  const double * ux; const double * uy; const double * uz; const double * t; const double * p;
  double * nx_; double * ny_; double * nz_; double * nt;          /* increment or new field */
  const double * sbx; const double * sby; const double * sbz; const double * sbt;   /* srcb */
  double * sax; double * say; double * saz; double * sat;         /* srca (next substep's b) */
  const double * dt;
  int rkstep;
  int final_here;   /* 1: no line solve follows, write u + delta with halos */
};

/* body with restrict-qualified pointers, so that the loads of each field section are free
 * to move ahead of the stores of the previous one */
__device__ __forceinline__ void predict_body(const PredictArgs & A,
    const double * __restrict__ ux, const double * __restrict__ uy, const double * __restrict__ uz,
    const double * __restrict__ t, const double * __restrict__ p,
    double * __restrict__ nxo, double * __restrict__ nyo, double * __restrict__ nzo, double * __restrict__ nto,
    const double * __restrict__ sbx, const double * __restrict__ sby, const double * __restrict__ sbz,
    const double * __restrict__ sbt,
    double * __restrict__ sax, double * __restrict__ say, double * __restrict__ saz, double * __restrict__ sat){
  const tc_geom_t & g = A.g;
  const tc_coef_t & c = A.c;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if(i > g.nx || j > g.ny) return;
  const long long o = tck_idx(g, i, j, k);
  const long long sy = g.px;
  const long long sz = g.plane;
  const double nu = c.nu;
  const double nu2 = 2. * nu;
  const double dt = *A.dt;
  const int s = A.rkstep;
  const double ca = c_rka[s] * dt;
  const double cb = c_rkb[s] * dt;
  const double cg = (c_rka[s] + c_rkb[s]) * dt;
  const bool ix = A.f.implicit[0], iy = A.f.implicit[1], iz = A.f.implicit[2];

  /* ---------------- ux, faces i = 2..nx (predict/ux.c) ---------------- */
  if(i >= 2){
    double a = 0., gg = 0.;
    const double u0 = ux[o], uxm = ux[o - 1], uxp = ux[o + 1];
    { /* advection_x :92-124 */
      const double ux_xm = c.axf[i - 1] * uxm + c.axf[i] * u0;
      const double ux_xp = c.axf[i] * u0 + c.axf[i + 1] * uxp;
      const double l = - 0.5 * ux_xm;
      const double u = + 0.5 * ux_xp;
      const double cc = - l - u;
      a -= c.ijdxf[i] * (+ l * uxm + cc * u0 + u * uxp);
    }
    { /* advection_y :126-159 */
      const double uy_ym = c.ayc[i - 1] * uy[o - 1] + c.ayc[i] * uy[o];
      const double uy_yp = c.ayc[i - 1] * uy[o - 1 + sy] + c.ayc[i] * uy[o + sy];
      const double l = - 0.5 * uy_ym;
      const double u = + 0.5 * uy_yp;
      const double cc = - l - u;
      a -= c.ijdxf[i] * (+ l * ux[o - sy] + cc * u0 + u * ux[o + sy]);
    }
    { /* advection_z :161-192 */
      const double uz_zm = c.azc[i - 1] * uz[o - 1] + c.azc[i] * uz[o];
      const double uz_zp = c.azc[i - 1] * uz[o - 1 + sz] + c.azc[i] * uz[o + sz];
      const double l = - 0.5 * uz_zm;
      const double u = + 0.5 * uz_zp;
      const double cc = - l - u;
      a -= c.ijdxf[i] * (+ l * ux[o - sz] + cc * u0 + u * ux[o + sz]);
    }
    { /* advection_1 :194-226 (centrifugal) */
      const double uy_xm = + 0.5 * uy[o - 1] + 0.5 * uy[o - 1 + sy];
      const double uy_xp = + 0.5 * uy[o] + 0.5 * uy[o + sy];
      a += c.ijdxf[i] * (+ c.d1m[i] * uy_xm * uy_xm + c.d1p[i] * uy_xp * uy_xp);
    }
    { /* diffusion_x :228-248 */
      const double v = nu2 * (+ c.lxu_l[i] * uxm + c.lxu_c[i] * u0 + c.lxu_u[i] * uxp);
      if(ix) gg += v; else a += v;
    }
    { /* diffusion_y0 :250-269 */
      const double v = nu * (+ c.lyu_l[i] * ux[o - sy] + c.lyu_c[i] * u0 + c.lyu_u[i] * ux[o + sy]);
      if(iy) gg += v; else a += v;
    }
    { /* diffusion_y1 :271-297; lyx1 (lyx.c:56-95) and lxy (lxy.c:19-49) at j and j+1 */
      const double lyx_ym = c.l1f[i] * (+ 0.5 * uy[o - 1] + 0.5 * uy[o]);
      const double lyx_yp = c.l1f[i] * (+ 0.5 * uy[o - 1 + sy] + 0.5 * uy[o + sy]);
      const double lxy_ym = c.ihxf[i] * (- uy[o - 1] + uy[o]);
      const double lxy_yp = c.ihxf[i] * (- uy[o - 1 + sy] + uy[o + sy]);
      a += c.nujf[i] * (- c.jhyf[i] * (lyx_ym + lxy_ym) + c.jhyf[i] * (lyx_yp + lxy_yp));
    }
    { /* diffusion_z0 :299-318 */
      const double v = nu * (+ c.lapz_l * ux[o - sz] + c.lapz_c * u0 + c.lapz_u * ux[o + sz]);
      if(iz) gg += v; else a += v;
    }
    { /* diffusion_z1 :320-342; lxz (lxz.c:10-44) at k and k+1 */
      const double lxz_zm = c.ihxf[i] * (- uz[o - 1] + uz[o]);
      const double lxz_zp = c.ihxf[i] * (- uz[o - 1 + sz] + uz[o + sz]);
      a += c.nujf[i] * (- c.jhzf[i] * lxz_zm + c.jhzf[i] * lxz_zp);
    }
    { /* diffusion_a :344-375; lyy0 (lyy.c:24-54) + lyy1 (lyy.c:56-95) at i-1 and i */
      const double lyy0_m = c.ihyc[i - 1] * (- uy[o - 1] + uy[o - 1 + sy]);
      const double lyy1_m = c.l1c[i - 1] * (+ 0.5 * uxm + 0.5 * u0);
      const double lyy0_p = c.ihyc[i] * (- uy[o] + uy[o + sy]);
      const double lyy1_p = c.l1c[i] * (+ 0.5 * u0 + 0.5 * uxp);
      const double lyy_xm = lyy0_m + lyy1_m;
      const double lyy_xp = lyy0_p + lyy1_p;
      a -= c.n2jf[i] * (+ c.d1m[i] * lyy_xm + c.d1p[i] * lyy_xp);
    }
    /* pressure :377-394 */
    gg -= c.ihxf[i] * (- p[o - 1] + p[o]);
    double b = 0.;
    if(sbx) b = sbx[o];
    const double delta = (sbx ? + ca * a + cb * b : + ca * a) + cg * gg;
    if(sax) sax[o] = a;
    if(A.final_here) tck_store_halo(nxo, g, i, j, k, u0 + delta, false);
    else nxo[o] = delta;
  }

  /* ---------------- uy, centres i = 1..nx (predict/uy.c) ---------------- */
  {
    double a = 0., gg = 0.;
    const double v0 = uy[o], vxm = uy[o - 1], vxp = uy[o + 1];
    const double vym = uy[o - sy], vyp = uy[o + sy];
    const double vzm = uy[o - sz], vzp = uy[o + sz];
    /* ux at (i, j-1), (i, j), (i+1, j-1), (i+1, j) */
    const double x00 = ux[o - sy], x01 = ux[o], x10 = ux[o + 1 - sy], x11 = ux[o + 1];
    { /* advection_x :93-126 */
      const double ux_xm = c.axf[i] * x00 + c.axf[i] * x01;
      const double ux_xp = c.axf[i + 1] * x10 + c.axf[i + 1] * x11;
      const double l = - 0.5 * ux_xm;
      const double u = + 0.5 * ux_xp;
      const double cc = - l - u;
      a -= c.ijdxc[i] * (+ l * vxm + cc * v0 + u * vxp);
    }
    { /* advection_y :128-156 */
      const double uy_ym = c.ayc[i] * vym + c.ayc[i] * v0;
      const double uy_yp = c.ayc[i] * v0 + c.ayc[i] * vyp;
      const double l = - 0.5 * uy_ym;
      const double u = + 0.5 * uy_yp;
      const double cc = - l - u;
      a -= c.ijdxc[i] * (+ l * vym + cc * v0 + u * vyp);
    }
    { /* advection_z :158-186 */
      const double uz_zm = c.azc[i] * uz[o - sy] + c.azc[i] * uz[o];
      const double uz_zp = c.azc[i] * uz[o - sy + sz] + c.azc[i] * uz[o + sz];
      const double l = - 0.5 * uz_zm;
      const double u = + 0.5 * uz_zp;
      const double cc = - l - u;
      a -= c.ijdxc[i] * (+ l * vzm + cc * v0 + u * vzp);
    }
    { /* advection_1 :188-223 (Coriolis) */
      const double uy_ym = + 0.5 * vym + 0.5 * v0;
      const double uy_yp = + 0.5 * v0 + 0.5 * vyp;
      const double ux_ym = + 0.5 * x00 + 0.5 * x10;
      const double ux_yp = + 0.5 * x01 + 0.5 * x11;
      a -= c.ijdxc[i] * (+ c.hdc[i] * uy_ym * ux_ym + c.hdc[i] * uy_yp * ux_yp);
    }
    { /* diffusion_x0 :225-244 */
      const double v = nu * (+ c.lxc_l[i] * vxm + c.lxc_c[i] * v0 + c.lxc_u[i] * vxp);
      if(ix) gg += v; else a += v;
    }
    /* lyx0 (lyx.c:24-54), lyx1 (:56-95) and lxy (lxy.c:19-49) at faces i and i+1 */
    const double lyx0_m = c.ihyf[i] * (- x00 + x01);
    const double lyx0_p = c.ihyf[i + 1] * (- x10 + x11);
    const double lyx1_m = c.l1f[i] * (+ (1 == i ? 0. : 0.5) * vxm + 0.5 * v0);
    const double lyx1_p = c.l1f[i + 1] * (+ 0.5 * v0 + (g.nx == i ? 0. : 0.5) * vxp);
    const double lyx_xm = lyx0_m + lyx1_m;
    const double lyx_xp = lyx0_p + lyx1_p;
    { /* diffusion_x1 :246-274 */
      a += c.nujc[i] * (- c.jhxf[i] * lyx_xm + c.jhxf[i + 1] * lyx_xp);
    }
    { /* diffusion_y0 :276-296 */
      const double v = nu2 * (+ c.lyc_l[i] * vym + c.lyc_c[i] * v0 + c.lyc_u[i] * vyp);
      if(iy) gg += v; else a += v;
    }
    { /* diffusion_y1 :298-320; lyy1 (lyy.c:56-95) at j-1 and j */
      const double lyy1_m = c.l1c[i] * (+ 0.5 * x00 + 0.5 * x10);
      const double lyy1_p = c.l1c[i] * (+ 0.5 * x01 + 0.5 * x11);
      a += c.n2jc[i] * (- c.jhyc[i] * lyy1_m + c.jhyc[i] * lyy1_p);
    }
    { /* diffusion_z0 :322-341 */
      const double v = nu * (+ c.lapz_l * vzm + c.lapz_c * v0 + c.lapz_u * vzp);
      if(iz) gg += v; else a += v;
    }
    { /* diffusion_z1 :343-365; lyz (lyz.c:10-44) at k and k+1 */
      const double lyz_zm = c.ihyc[i] * (- uz[o - sy] + uz[o]);
      const double lyz_zp = c.ihyc[i] * (- uz[o - sy + sz] + uz[o + sz]);
      a += c.nujc[i] * (- c.jhzc[i] * lyz_zm + c.jhzc[i] * lyz_zp);
    }
    { /* diffusion_a :367-400 */
      const double lxy_xm = c.ihxf[i] * (- vxm + v0);
      const double lxy_xp = c.ihxf[i + 1] * (- v0 + vxp);
      a += c.nujc[i] * (+ c.dam[i] * (lyx_xm + lxy_xm) + c.dap[i] * (lyx_xp + lxy_xp));
    }
    /* pressure :402-419 */
    gg -= c.ihyc[i] * (- p[o - sy] + p[o]);
    double b = 0.;
    if(sby) b = sby[o];
    const double delta = (sby ? + ca * a + cb * b : + ca * a) + cg * gg;
    if(say) say[o] = a;
    if(A.final_here) tck_store_halo(nyo, g, i, j, k, v0 + delta, false);
    else nyo[o] = delta;
  }

  /* ---------------- uz, centres (predict/uz.c) ---------------- */
  {
    double a = 0., gg = 0.;
    const double w0 = uz[o], wxm = uz[o - 1], wxp = uz[o + 1];
    const double wym = uz[o - sy], wyp = uz[o + sy];
    const double wzm = uz[o - sz], wzp = uz[o + sz];
    /* ux at (i, k-1), (i, k), (i+1, k-1), (i+1, k) */
    const double x00 = ux[o - sz], x01 = ux[o], x10 = ux[o + 1 - sz], x11 = ux[o + 1];
    /* uy at (j, k-1), (j, k), (j+1, k-1), (j+1, k) */
    const double y00 = uy[o - sz], y01 = uy[o], y10 = uy[o + sy - sz], y11 = uy[o + sy];
    { /* advection_x :88-121 */
      const double ux_xm = c.axf[i] * x00 + c.axf[i] * x01;
      const double ux_xp = c.axf[i + 1] * x10 + c.axf[i + 1] * x11;
      const double l = - 0.5 * ux_xm;
      const double u = + 0.5 * ux_xp;
      const double cc = - l - u;
      a -= c.ijdxc[i] * (+ l * wxm + cc * w0 + u * wxp);
    }
    { /* advection_y :123-152 */
      const double uy_ym = c.ayc[i] * y00 + c.ayc[i] * y01;
      const double uy_yp = c.ayc[i] * y10 + c.ayc[i] * y11;
      const double l = - 0.5 * uy_ym;
      const double u = + 0.5 * uy_yp;
      const double cc = - l - u;
      a -= c.ijdxc[i] * (+ l * wym + cc * w0 + u * wyp);
    }
    { /* advection_z :154-181 */
      const double uz_zm = c.azc[i] * wzm + c.azc[i] * w0;
      const double uz_zp = c.azc[i] * w0 + c.azc[i] * wzp;
      const double l = - 0.5 * uz_zm;
      const double u = + 0.5 * uz_zp;
      const double cc = - l - u;
      a -= c.ijdxc[i] * (+ l * wzm + cc * w0 + u * wzp);
    }
    { /* diffusion_x0 :183-202 */
      const double v = nu * (+ c.lxc_l[i] * wxm + c.lxc_c[i] * w0 + c.lxc_u[i] * wxp);
      if(ix) gg += v; else a += v;
    }
    { /* diffusion_x1 :204-231; lzx (lzx.c:9-42) at faces i and i+1 */
      const double lzx_xm = c.ihz * (- x00 + x01);
      const double lzx_xp = c.ihz * (- x10 + x11);
      a += c.nujc[i] * (- c.jhxf[i] * lzx_xm + c.jhxf[i + 1] * lzx_xp);
    }
    { /* diffusion_y0 :233-252 */
      const double v = nu * (+ c.lyc_l[i] * wym + c.lyc_c[i] * w0 + c.lyc_u[i] * wyp);
      if(iy) gg += v; else a += v;
    }
    { /* diffusion_y1 :254-277; lzy (lzy.c:9-42) at j and j+1 */
      const double lzy_ym = c.ihz * (- y00 + y01);
      const double lzy_yp = c.ihz * (- y10 + y11);
      a += c.nujc[i] * (- c.jhyc[i] * lzy_ym + c.jhyc[i] * lzy_yp);
    }
    { /* diffusion_z :279-299 */
      const double v = nu2 * (+ c.lapz_l * wzm + c.lapz_c * w0 + c.lapz_u * wzp);
      if(iz) gg += v; else a += v;
    }
    /* pressure :301-318 */
    gg -= c.ihz * (- p[o - sz] + p[o]);
    double b = 0.;
    if(sbz) b = sbz[o];
    const double delta = (sbz ? + ca * a + cb * b : + ca * a) + cg * gg;
    if(saz) saz[o] = a;
    if(A.final_here) tck_store_halo(nzo, g, i, j, k, w0 + delta, false);
    else nzo[o] = delta;
  }

  /* ---------------- T, centres (predict/t.c) ---------------- */
  if(A.f.scalar){
    const double kappa = c.kappa;
    double a = 0., gg = 0.;
    const double t0 = t[o], txm = t[o - 1], txp = t[o + 1];
    const double tym = t[o - sy], typ = t[o + sy];
    const double tzm = t[o - sz], tzp = t[o + sz];
    { /* advection_x :88-119 */
      const double ux_xm = c.jhxf[i] * ux[o];
      const double ux_xp = c.jhxf[i + 1] * ux[o + 1];
      const double l = - 0.5 * ux_xm;
      const double u = + 0.5 * ux_xp;
      const double cc = - l - u;
      a -= c.ijdxc[i] * (+ l * txm + cc * t0 + u * txp);
    }
    { /* advection_y :121-148 */
      const double uy_ym = c.jhyc[i] * uy[o];
      const double uy_yp = c.jhyc[i] * uy[o + sy];
      const double l = - 0.5 * uy_ym;
      const double u = + 0.5 * uy_yp;
      const double cc = - l - u;
      a -= c.ijdxc[i] * (+ l * tym + cc * t0 + u * typ);
    }
    { /* advection_z :150-176 */
      const double uz_zm = c.jhzc[i] * uz[o];
      const double uz_zp = c.jhzc[i] * uz[o + sz];
      const double l = - 0.5 * uz_zm;
      const double u = + 0.5 * uz_zp;
      const double cc = - l - u;
      a -= c.ijdxc[i] * (+ l * tzm + cc * t0 + u * tzp);
    }
    { /* diffusion_x/y/z :178-239 */
      const double vx = kappa * (+ c.lxc_l[i] * txm + c.lxc_c[i] * t0 + c.lxc_u[i] * txp);
      if(ix) gg += vx; else a += vx;
      const double vy = kappa * (+ c.lyc_l[i] * tym + c.lyc_c[i] * t0 + c.lyc_u[i] * typ);
      if(iy) gg += vy; else a += vy;
      const double vz = kappa * (+ c.lapz_l * tzm + c.lapz_c * t0 + c.lapz_u * tzp);
      if(iz) gg += vz; else a += vz;
    }
    double b = 0.;
    if(sbt) b = sbt[o];
    const double delta = (sbt ? + ca * a + cb * b : + ca * a) + cg * gg;
    if(sat) sat[o] = a;
    if(A.final_here) tck_store_halo(nto, g, i, j, k, t0 + delta, false);
    else nto[o] = delta;
  }
}

__global__ void __launch_bounds__(TCK_BLOCK) tck_predict(const PredictArgs A){
  predict_body(A, A.ux, A.uy, A.uz, A.t, A.p, A.nx_, A.ny_, A.nz_, A.nt, A.sbx, A.sby, A.sbz, A.sbt,
      A.sax, A.say, A.saz, A.sat);
}

/* ------------------------------------------------------------------------------------- */

struct CorrectArgs {
  tc_geom_t g;
  tc_coef_t c;
  tc_flags_t f;
  double * ux; double * uy; double * uz; double * p;
  const double * psi;
  const double * dt;
  int rkstep;
  int mode;                     /* bit 0: velocity correction, bit 1: pressure update */
  unsigned long long * dtmin;   /* non-null: CFL minimum of the corrected field */
};

__device__ __forceinline__ void correct_body(const CorrectArgs & A, double * __restrict__ ux,
    double * __restrict__ uy, double * __restrict__ uz, double * __restrict__ pp,
    const double * __restrict__ psi){
  const tc_geom_t & g = A.g;
  const tc_coef_t & c = A.c;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  const bool inside = (i <= g.nx && j <= g.ny);
  double lmin = 1.;
  if(inside){
    const long long o = tck_idx(g, i, j, k);
    const long long sy = g.px;
    const long long sz = g.plane;
    const double dt = *A.dt;
    const int s = A.rkstep;
    const double gamma = c_rka[s] + c_rkb[s];
    const double prefactor = gamma * dt;                /* correct/main.c:22-23 */
    const double small = 1.e-8;
    const double ps0 = psi[o];
    if(A.mode & 1){
    if(i >= 2){ /* correct/ux.c:25-52 */
      const double hx = c.hxxf[i];
      const double v = ux[o] - prefactor / hx * (- psi[o - 1] + ps0);
      tck_store_halo(ux, g, i, j, k, v, false);
      lmin = fmin(lmin, hx / (fabs(v) + small));
    }
    { /* correct/uy.c:25-52 */
      const double hy = c.hyxc[i];
      const double v = uy[o] - prefactor / hy * (- psi[o - sy] + ps0);
      tck_store_halo(uy, g, i, j, k, v, false);
      lmin = fmin(lmin, hy / (fabs(v) + small));
    }
    { /* correct/uz.c:15-44 */
      const double v = uz[o] - prefactor / c.hz * (- psi[o - sz] + ps0);
      tck_store_halo(uz, g, i, j, k, v, false);
      lmin = fmin(lmin, c.hz / (fabs(v) + small));
    }
    }
    if(A.mode & 2){ /* update_pressure.c:146-173 */
      double v = pp[o] + ps0;
      const double pf = 0.5 * gamma * dt * c.nu;
      if(A.f.implicit[0]){
        const double l = c.plx_l[i];
        const double u = c.plx_u[i];
        const double cc = - l - u;
        v -= 2. * pf / c.jdxc[i] * (+ l * psi[o - 1] + cc * ps0 + u * psi[o + 1]);
      }
      if(A.f.implicit[1]){
        const double l = c.ply[i];
        const double u = c.ply[i];
        const double cc = - l - u;
        v -= 2. * pf / c.jdxc[i] * (+ l * psi[o - sy] + cc * ps0 + u * psi[o + sy]);
      }
      if(A.f.implicit[2]){
        const double l = c.plz[i];
        const double u = c.plz[i];
        const double cc = - l - u;
        v -= 2. * pf / c.jdxc[i] * (+ l * psi[o - sz] + cc * ps0 + u * psi[o + sz]);
      }
      tck_store_halo(pp, g, i, j, k, v, true);
    }
  }
  if(A.dtmin){
    /* CFL bound of decide_dt_adv (decide_dt.c:32-80) on the corrected velocity */
    __shared__ double red[TCK_BLOCK / TCK_WAVE];
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    double m = tck_wave_min(lmin);
    if((tid % TCK_WAVE) == 0) red[tid / TCK_WAVE] = m;
    __syncthreads();
    if(tid == 0){
      const int nw = (blockDim.x * blockDim.y + TCK_WAVE - 1) / TCK_WAVE;
      for(int w = 1; w < nw; w++) m = fmin(m, red[w]);
      atomicMin(A.dtmin, tck_bits(m));
    }
  }
}

__global__ void __launch_bounds__(TCK_BLOCK) tck_correct(const CorrectArgs A){
  correct_body(A, A.ux, A.uy, A.uz, A.p, A.psi);
}

/* decide_dt_adv on the current field (decide_dt.c:32-80), for the first step and after
 * fields are replaced from the host */
struct DtArgs {
  tc_geom_t g;
  tc_coef_t c;
  const double * ux; const double * uy; const double * uz;
  unsigned long long * dtmin;
};

__global__ void __launch_bounds__(TCK_BLOCK) tck_dt_reduce(const DtArgs A){
  const tc_geom_t & g = A.g;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  double lmin = 1.;
  const double small = 1.e-8;
  if(i <= g.nx && j <= g.ny){
    const long long o = tck_idx(g, i, j, k);
    if(i >= 2) lmin = fmin(lmin, A.c.hxxf[i] / (fabs(A.ux[o]) + small));
    lmin = fmin(lmin, A.c.hyxc[i] / (fabs(A.uy[o]) + small));
    lmin = fmin(lmin, A.c.hz / (fabs(A.uz[o]) + small));
  }
  __shared__ double red[TCK_BLOCK / TCK_WAVE];
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  double m = tck_wave_min(lmin);
  if((tid % TCK_WAVE) == 0) red[tid / TCK_WAVE] = m;
  __syncthreads();
  if(tid == 0){
    const int nw = (blockDim.x * blockDim.y + TCK_WAVE - 1) / TCK_WAVE;
    for(int w = 1; w < nw; w++) m = fmin(m, red[w]);
    atomicMin(A.dtmin, tck_bits(m));
  }
}

/* decide_dt.c:125-160: dt = coef_adv * min(1, CFL) then min with the explicit diffusive
 * limits (precomputed on the host, decide_dt.c:89-115); re-arms the CFL accumulator */
__global__ void tck_dt_finalise(unsigned long long * dtmin, double * dt, double coef_adv,
    double dif_limit, double dt_fixed){
  if(threadIdx.x != 0 || blockIdx.x != 0) return;
  if(dt_fixed > 0.){
    *dt = dt_fixed;
  }else{
    double v = __longlong_as_double((long long)*dtmin);
    v *= coef_adv;
    *dt = fmin(v, dif_limit);
  }
  *dtmin = tck_bits(1.);
}

__global__ void tck_dt_arm(unsigned long long * dtmin){
  if(threadIdx.x == 0 && blockIdx.x == 0) *dtmin = tck_bits(1.);
}

/* ------------------------------------------------------------------------------------- */
extern "C" {

static dim3 cell_grid(const tc_geom_t & g, dim3 * blk){
  *blk = dim3(64, 4, 1);
  return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
}

int tck_launch_predict(const tc_geom_t * g, const tc_coef_t * c, const tc_flags_t * f,
    const double * const cur[4], const double * p, double * const nxt[4],
    const double * const srcb[4], double * const srca[4], const double * dt, int rkstep,
    int final_here, hipStream_t stream){
  PredictArgs A;
  A.g = *g; A.c = *c; A.f = *f;
  A.ux = cur[0]; A.uy = cur[1]; A.uz = cur[2]; A.t = cur[3]; A.p = p;
  A.nx_ = nxt[0]; A.ny_ = nxt[1]; A.nz_ = nxt[2]; A.nt = nxt[3];
  A.sbx = srcb[0]; A.sby = srcb[1]; A.sbz = srcb[2]; A.sbt = srcb[3];
  A.sax = srca[0]; A.say = srca[1]; A.saz = srca[2]; A.sat = srca[3];
  A.dt = dt; A.rkstep = rkstep; A.final_here = final_here;
  dim3 blk;
  dim3 grd = cell_grid(*g, &blk);
  hipLaunchKernelGGL(tck_predict, grd, blk, 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_correct(const tc_geom_t * g, const tc_coef_t * c, const tc_flags_t * f,
    double * ux, double * uy, double * uz, double * p, const double * psi, const double * dt,
    int rkstep, int mode, unsigned long long * dtmin, hipStream_t stream){
  CorrectArgs A;
  A.mode = mode;
  A.g = *g; A.c = *c; A.f = *f;
  A.ux = ux; A.uy = uy; A.uz = uz; A.p = p; A.psi = psi;
  A.dt = dt; A.rkstep = rkstep; A.dtmin = dtmin;
  dim3 blk;
  dim3 grd = cell_grid(*g, &blk);
  hipLaunchKernelGGL(tck_correct, grd, blk, 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_dt(const tc_geom_t * g, const tc_coef_t * c, const double * ux, const double * uy,
    const double * uz, unsigned long long * dtmin, hipStream_t stream){
  DtArgs A;
  A.g = *g; A.c = *c; A.ux = ux; A.uy = uy; A.uz = uz; A.dtmin = dtmin;
  hipLaunchKernelGGL(tck_dt_arm, dim3(1), dim3(1), 0, stream, dtmin);
  TCK_CHECK_LAUNCH();
  dim3 blk;
  dim3 grd = cell_grid(*g, &blk);
  hipLaunchKernelGGL(tck_dt_reduce, grd, blk, 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_dt_finalise(unsigned long long * dtmin, double * dt, double coef_adv,
    double dif_limit, double dt_fixed, hipStream_t stream){
  hipLaunchKernelGGL(tck_dt_finalise, dim3(1), dim3(1), 0, stream, dtmin, dt, coef_adv, dif_limit, dt_fixed);
  TCK_CHECK_LAUNCH();
  return 0;
}

}
