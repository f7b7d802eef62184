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

/* Launch shape of the cell kernels: a block of 64 x ty threads covers an (i, j) tile and
 * marches kb planes in k, so the k-1 and k+1 planes a thread reads were read by the same block
 * one iteration earlier.  Blocks are numbered tile-major (i tile fastest, then j, then the k
 * chunk) and renumbered so that each of the 8 XCDs (which receive launch-order block ids
 * round-robin and keep their own L2) works through one contiguous eighth of the tiles: the
 * halo rows and planes a block shares with its neighbours then sit in the same L2. */
#define TCK_DTSLOTS 64      /* CFL minimum accumulators, one 128 B line apart */
#define TCK_DTSTRIDE 16

struct TileMap {
  int nbx, nby, nbz;   /* tiles in i, j and k chunks */
  int kb;              /* planes per block */
  long long total;
};

__device__ __forceinline__ void tile_of(const TileMap & T, int * bx, int * by, int * bz){
  long long t = blockIdx.x;
  if(T.total % TCK_NXCD == 0){
    const long long per = T.total / TCK_NXCD;
    t = (t % TCK_NXCD) * per + t / TCK_NXCD;
  }
  *bx = (int)(t % T.nbx);
  t /= T.nbx;
  *by = (int)(t % T.nby);
  *bz = (int)(t / T.nby);
}

struct PredictArgs {
  tc_geom_t g;
  tc_coef_t c;
  tc_flags_t f;
  const double * ux; const double * uy; const double * uz; const double * t; const double * p;
  double * nx_; double * ny_; double * nz_; double * nt;          /* new field (final_here) */
  /* buffers of the arithmetic type R of the instantiation (double, or float in the mixed
   * step): the increment, srcb and srca (next substep's b) */
  void * dx; void * dy; void * dz; void * dtt;
  const void * sbx; const void * sby; const void * sbz; const void * sbt;
  void * sax; void * say; void * saz; void * sat;
  /* learned corrections of the explicit right-hand sides (field layout, float), or null */
  const float * cx; const float * cy; const float * cz; const float * ct;
  const double * dt;
  int rkstep;
  int final_here;   /* 1: no line solve follows, write u + delta with halos */
  TileMap tm;
};

/* coefficient row t at index ii, from the block's LDS copy of rows 0..T_NPRED-1 over the
 * window ib .. ib + TCK_TW - 1 */
#define TCK_TW (TCK_WAVE + 2)
/* keeps the loads of one field's section from being hoisted above the previous section's
 * stores: shorter live ranges, fewer registers, higher occupancy */
#define TCK_SECTION_FENCE() asm volatile("" ::: "memory")
#define CT(t, ii) (smt[(t) * TW + (ii) - ib])

/* the FP64 value of a cell's own field: from the staged plane when it is staged in FP64,
 * else from the field in memory (the window of the mixed step stages FP32) */
template <typename FT>
__device__ __forceinline__ double own_value(const FT * staged, int of, const double * field, long long o){
  return sizeof(FT) == 8 ? double(staged[of]) : field[o];
}

/* body with restrict-qualified pointers, so that the loads of each field section are free
 * to move ahead of the stores of the previous one */
template <typename R, typename PT, typename FT, int TW, int CM, int IMP = -1>
__device__ __forceinline__ void predict_body(const PredictArgs & A,
    const FT * __restrict__ Fux_m, const FT * __restrict__ Fux_0, const FT * __restrict__ Fux_p,
    const FT * __restrict__ Fuy_m, const FT * __restrict__ Fuy_0, const FT * __restrict__ Fuy_p,
    const FT * __restrict__ Fuz_m, const FT * __restrict__ Fuz_0, const FT * __restrict__ Fuz_p,
    const FT * __restrict__ Ft_m, const FT * __restrict__ Ft_0, const FT * __restrict__ Ft_p,
    const PT * __restrict__ Fp_m, const PT * __restrict__ Fp_0, const int of, const int fsy,
    R * __restrict__ dxo, R * __restrict__ dyo, R * __restrict__ dzo, R * __restrict__ dto,
    double * __restrict__ nxo, double * __restrict__ nyo, double * __restrict__ nzo, double * __restrict__ nto,
    const R * __restrict__ sbx, const R * __restrict__ sby, const R * __restrict__ sbz,
    const R * __restrict__ sbt, const R * sbv,
    R * __restrict__ sax, R * __restrict__ say, R * __restrict__ saz, R * __restrict__ sat,
    const int i, const int j, const int k, const long long kb0, const R * smt, const int ib,
    const R dt, const float * cvv){
  const tc_geom_t & g = A.g;
  const tc_coef_t & c = A.c;
  if(i > g.nx || j > g.ny) return;
  /* the field pointers start at plane k-1 (kb0): offsets inside the 3-plane window fit in
   * 32 bits, so every load is a base + 32-bit offset */
  const int sy = g.px;
  const int sz = (int)g.plane;
  const int o = (i + TC_XOFF) + sy * j + sz;
  const R nu = R(c.nu);
  const R nu2 = R(2.) * nu;
  const int s = A.rkstep;
  const R ca = c_rka[s] * dt;
  const R cb = c_rkb[s] * dt;
  const R cg = (c_rka[s] + c_rkb[s]) * dt;
  /* IMP >= 0: the implicit directions as compile-time bits (a line solve follows, so no
   * final store); IMP < 0: from the flags */
  const bool ix = IMP >= 0 ? (IMP & 1) != 0 : A.f.implicit[0];
  const bool iy = IMP >= 0 ? (IMP & 2) != 0 : A.f.implicit[1];
  const bool iz = IMP >= 0 ? (IMP & 4) != 0 : A.f.implicit[2];
  const bool fin = IMP >= 0 ? false : A.final_here != 0;

  /* ---------------- ux, faces i = 2..nx (predict/ux.c) ---------------- */
  if(i >= 2){
    R a = R(0.), gg = R(0.);
    const R u0 = R(Fux_0[of]), uxm = R(Fux_0[of - 1]), uxp = R(Fux_0[of + 1]);
    { /* advection_x :92-124 */
      const R ux_xm = CT(T_AXF, i - 1) * uxm + CT(T_AXF, i) * u0;
      const R ux_xp = CT(T_AXF, i) * u0 + CT(T_AXF, i + 1) * uxp;
      const R l = - R(0.5) * ux_xm;
      const R u = + R(0.5) * ux_xp;
      const R cc = - l - u;
      a -= CT(T_IJDXF, i) * (+ l * uxm + cc * u0 + u * uxp);
    }
    { /* advection_y :126-159 */
      const R uy_ym = CT(T_AYC, i - 1) * R(Fuy_0[of - 1]) + CT(T_AYC, i) * R(Fuy_0[of]);
      const R uy_yp = CT(T_AYC, i - 1) * R(Fuy_0[of - 1 + fsy]) + CT(T_AYC, i) * R(Fuy_0[of + fsy]);
      const R l = - R(0.5) * uy_ym;
      const R u = + R(0.5) * uy_yp;
      const R cc = - l - u;
      a -= CT(T_IJDXF, i) * (+ l * R(Fux_0[of - fsy]) + cc * u0 + u * R(Fux_0[of + fsy]));
    }
    { /* advection_z :161-192 */
      const R uz_zm = CT(T_AZC, i - 1) * R(Fuz_0[of - 1]) + CT(T_AZC, i) * R(Fuz_0[of]);
      const R uz_zp = CT(T_AZC, i - 1) * R(Fuz_p[of - 1]) + CT(T_AZC, i) * R(Fuz_p[of]);
      const R l = - R(0.5) * uz_zm;
      const R u = + R(0.5) * uz_zp;
      const R cc = - l - u;
      a -= CT(T_IJDXF, i) * (+ l * R(Fux_m[of]) + cc * u0 + u * R(Fux_p[of]));
    }
    { /* advection_1 :194-226 (centrifugal) */
      const R uy_xm = + R(0.5) * R(Fuy_0[of - 1]) + R(0.5) * R(Fuy_0[of - 1 + fsy]);
      const R uy_xp = + R(0.5) * R(Fuy_0[of]) + R(0.5) * R(Fuy_0[of + fsy]);
      a += CT(T_IJDXF, i) * (+ CT(T_D1M, i) * uy_xm * uy_xm + CT(T_D1P, i) * uy_xp * uy_xp);
    }
    { /* diffusion_x :228-248 */
      const R v = nu2 * (+ CT(T_LXU_L, i) * uxm + CT(T_LXU_C, i) * u0 + CT(T_LXU_U, i) * uxp);
      if(ix) gg += v; else a += v;
    }
    { /* diffusion_y0 :250-269 */
      const R v = nu * (+ CT(T_LYU_L, i) * R(Fux_0[of - fsy]) + CT(T_LYU_C, i) * u0 + CT(T_LYU_U, i) * R(Fux_0[of + fsy]));
      if(iy) gg += v; else a += v;
    }
    { /* diffusion_y1 :271-297; lyx1 (lyx.c:56-95) and lxy (lxy.c:19-49) at j and j+1 */
      const R lyx_ym = CT(T_L1F, i) * (+ R(0.5) * R(Fuy_0[of - 1]) + R(0.5) * R(Fuy_0[of]));
      const R lyx_yp = CT(T_L1F, i) * (+ R(0.5) * R(Fuy_0[of - 1 + fsy]) + R(0.5) * R(Fuy_0[of + fsy]));
      const R lxy_ym = CT(T_IHXF, i) * (- R(Fuy_0[of - 1]) + R(Fuy_0[of]));
      const R lxy_yp = CT(T_IHXF, i) * (- R(Fuy_0[of - 1 + fsy]) + R(Fuy_0[of + fsy]));
      a += CT(T_NUJF, i) * (- CT(T_JHYF, i) * (lyx_ym + lxy_ym) + CT(T_JHYF, i) * (lyx_yp + lxy_yp));
    }
    { /* diffusion_z0 :299-318 */
      const R v = nu * (+ R(c.lapz_l) * R(Fux_m[of]) + R(c.lapz_c) * u0 + R(c.lapz_u) * R(Fux_p[of]));
      if(iz) gg += v; else a += v;
    }
    { /* diffusion_z1 :320-342; lxz (lxz.c:10-44) at k and k+1 */
      const R lxz_zm = CT(T_IHXF, i) * (- R(Fuz_0[of - 1]) + R(Fuz_0[of]));
      const R lxz_zp = CT(T_IHXF, i) * (- R(Fuz_p[of - 1]) + R(Fuz_p[of]));
      a += CT(T_NUJF, i) * (- CT(T_JHZF, i) * lxz_zm + CT(T_JHZF, i) * lxz_zp);
    }
    { /* diffusion_a :344-375; lyy0 (lyy.c:24-54) + lyy1 (lyy.c:56-95) at i-1 and i */
      const R lyy0_m = CT(T_IHYC, i - 1) * (- R(Fuy_0[of - 1]) + R(Fuy_0[of - 1 + fsy]));
      const R lyy1_m = CT(T_L1C, i - 1) * (+ R(0.5) * uxm + R(0.5) * u0);
      const R lyy0_p = CT(T_IHYC, i) * (- R(Fuy_0[of]) + R(Fuy_0[of + fsy]));
      const R lyy1_p = CT(T_L1C, i) * (+ R(0.5) * u0 + R(0.5) * uxp);
      const R lyy_xm = lyy0_m + lyy1_m;
      const R lyy_xp = lyy0_p + lyy1_p;
      a -= CT(T_N2JF, i) * (+ CT(T_D1M, i) * lyy_xm + CT(T_D1P, i) * lyy_xp);
    }
    /* pressure :377-394 */
    gg -= CT(T_IHXF, i) * (- R(Fp_0[of - 1]) + R(Fp_0[of]));
    if(CM > 0) a += R(cvv[0]);
    else if(CM < 0 && A.cx) a += R(A.cx[kb0 + o]);
    R b = R(0.);
    if(sbx) b = sbv ? sbv[0] : sbx[o];
    const R delta = (sbx ? + ca * a + cb * b : + ca * a) + cg * gg;
    if(sax) tck_st(&sax[o], a);
    if(fin) tck_store_halo(nxo - kb0, g, i, j, k, own_value(Fux_0, of, A.ux, kb0 + o) + (double)delta, false);
    else tck_st(&dxo[o], delta);
  }

  TCK_SECTION_FENCE();
  /* ---------------- uy, centres i = 1..nx (predict/uy.c) ---------------- */
  {
    R a = R(0.), gg = R(0.);
    const R v0 = R(Fuy_0[of]), vxm = R(Fuy_0[of - 1]), vxp = R(Fuy_0[of + 1]);
    const R vym = R(Fuy_0[of - fsy]), vyp = R(Fuy_0[of + fsy]);
    const R vzm = R(Fuy_m[of]), vzp = R(Fuy_p[of]);
    /* ux at (i, j-1), (i, j), (i+1, j-1), (i+1, j) */
    const R x00 = R(Fux_0[of - fsy]), x01 = R(Fux_0[of]), x10 = R(Fux_0[of + 1 - fsy]), x11 = R(Fux_0[of + 1]);
    { /* advection_x :93-126 */
      const R ux_xm = CT(T_AXF, i) * x00 + CT(T_AXF, i) * x01;
      const R ux_xp = CT(T_AXF, i + 1) * x10 + CT(T_AXF, i + 1) * x11;
      const R l = - R(0.5) * ux_xm;
      const R u = + R(0.5) * ux_xp;
      const R cc = - l - u;
      a -= CT(T_IJDXC, i) * (+ l * vxm + cc * v0 + u * vxp);
    }
    { /* advection_y :128-156 */
      const R uy_ym = CT(T_AYC, i) * vym + CT(T_AYC, i) * v0;
      const R uy_yp = CT(T_AYC, i) * v0 + CT(T_AYC, i) * vyp;
      const R l = - R(0.5) * uy_ym;
      const R u = + R(0.5) * uy_yp;
      const R cc = - l - u;
      a -= CT(T_IJDXC, i) * (+ l * vym + cc * v0 + u * vyp);
    }
    { /* advection_z :158-186 */
      const R uz_zm = CT(T_AZC, i) * R(Fuz_0[of - fsy]) + CT(T_AZC, i) * R(Fuz_0[of]);
      const R uz_zp = CT(T_AZC, i) * R(Fuz_p[of - fsy]) + CT(T_AZC, i) * R(Fuz_p[of]);
      const R l = - R(0.5) * uz_zm;
      const R u = + R(0.5) * uz_zp;
      const R cc = - l - u;
      a -= CT(T_IJDXC, i) * (+ l * vzm + cc * v0 + u * vzp);
    }
    { /* advection_1 :188-223 (Coriolis) */
      const R uy_ym = + R(0.5) * vym + R(0.5) * v0;
      const R uy_yp = + R(0.5) * v0 + R(0.5) * vyp;
      const R ux_ym = + R(0.5) * x00 + R(0.5) * x10;
      const R ux_yp = + R(0.5) * x01 + R(0.5) * x11;
      a -= CT(T_IJDXC, i) * (+ CT(T_HDC, i) * uy_ym * ux_ym + CT(T_HDC, i) * uy_yp * ux_yp);
    }
    { /* diffusion_x0 :225-244 */
      const R v = nu * (+ CT(T_LXC_L, i) * vxm + CT(T_LXC_C, i) * v0 + CT(T_LXC_U, i) * vxp);
      if(ix) gg += v; else a += v;
    }
    /* lyx0 (lyx.c:24-54), lyx1 (:56-95) and lxy (lxy.c:19-49) at faces i and i+1 */
    const R lyx0_m = CT(T_IHYF, i) * (- x00 + x01);
    const R lyx0_p = CT(T_IHYF, i + 1) * (- x10 + x11);
    const R lyx1_m = CT(T_L1F, i) * (+ (1 == i ? R(0.) : R(0.5)) * vxm + R(0.5) * v0);
    const R lyx1_p = CT(T_L1F, i + 1) * (+ R(0.5) * v0 + (g.nx == i ? R(0.) : R(0.5)) * vxp);
    const R lyx_xm = lyx0_m + lyx1_m;
    const R lyx_xp = lyx0_p + lyx1_p;
    { /* diffusion_x1 :246-274 */
      a += CT(T_NUJC, i) * (- CT(T_JHXF, i) * lyx_xm + CT(T_JHXF, i + 1) * lyx_xp);
    }
    { /* diffusion_y0 :276-296 */
      const R v = nu2 * (+ CT(T_LYC_L, i) * vym + CT(T_LYC_C, i) * v0 + CT(T_LYC_U, i) * vyp);
      if(iy) gg += v; else a += v;
    }
    { /* diffusion_y1 :298-320; lyy1 (lyy.c:56-95) at j-1 and j */
      const R lyy1_m = CT(T_L1C, i) * (+ R(0.5) * x00 + R(0.5) * x10);
      const R lyy1_p = CT(T_L1C, i) * (+ R(0.5) * x01 + R(0.5) * x11);
      a += CT(T_N2JC, i) * (- CT(T_JHYC, i) * lyy1_m + CT(T_JHYC, i) * lyy1_p);
    }
    { /* diffusion_z0 :322-341 */
      const R v = nu * (+ R(c.lapz_l) * vzm + R(c.lapz_c) * v0 + R(c.lapz_u) * vzp);
      if(iz) gg += v; else a += v;
    }
    { /* diffusion_z1 :343-365; lyz (lyz.c:10-44) at k and k+1 */
      const R lyz_zm = CT(T_IHYC, i) * (- R(Fuz_0[of - fsy]) + R(Fuz_0[of]));
      const R lyz_zp = CT(T_IHYC, i) * (- R(Fuz_p[of - fsy]) + R(Fuz_p[of]));
      a += CT(T_NUJC, i) * (- CT(T_JHZC, i) * lyz_zm + CT(T_JHZC, i) * lyz_zp);
    }
    { /* diffusion_a :367-400 */
      const R lxy_xm = CT(T_IHXF, i) * (- vxm + v0);
      const R lxy_xp = CT(T_IHXF, i + 1) * (- v0 + vxp);
      a += CT(T_NUJC, i) * (+ CT(T_DAM, i) * (lyx_xm + lxy_xm) + CT(T_DAP, i) * (lyx_xp + lxy_xp));
    }
    /* pressure :402-419 */
    gg -= CT(T_IHYC, i) * (- R(Fp_0[of - fsy]) + R(Fp_0[of]));
    if(CM > 0) a += R(cvv[1]);
    else if(CM < 0 && A.cy) a += R(A.cy[kb0 + o]);
    R b = R(0.);
    if(sby) b = sbv ? sbv[1] : sby[o];
    const R delta = (sby ? + ca * a + cb * b : + ca * a) + cg * gg;
    if(say) tck_st(&say[o], a);
    if(fin) tck_store_halo(nyo - kb0, g, i, j, k, own_value(Fuy_0, of, A.uy, kb0 + o) + (double)delta, false);
    else tck_st(&dyo[o], delta);
  }

  TCK_SECTION_FENCE();
  /* ---------------- uz, centres (predict/uz.c) ---------------- */
  {
    R a = R(0.), gg = R(0.);
    const R w0 = R(Fuz_0[of]), wxm = R(Fuz_0[of - 1]), wxp = R(Fuz_0[of + 1]);
    const R wym = R(Fuz_0[of - fsy]), wyp = R(Fuz_0[of + fsy]);
    const R wzm = R(Fuz_m[of]), wzp = R(Fuz_p[of]);
    /* ux at (i, k-1), (i, k), (i+1, k-1), (i+1, k) */
    const R x00 = R(Fux_m[of]), x01 = R(Fux_0[of]), x10 = R(Fux_m[of + 1]), x11 = R(Fux_0[of + 1]);
    /* uy at (j, k-1), (j, k), (j+1, k-1), (j+1, k) */
    const R y00 = R(Fuy_m[of]), y01 = R(Fuy_0[of]), y10 = R(Fuy_m[of + fsy]), y11 = R(Fuy_0[of + fsy]);
    { /* advection_x :88-121 */
      const R ux_xm = CT(T_AXF, i) * x00 + CT(T_AXF, i) * x01;
      const R ux_xp = CT(T_AXF, i + 1) * x10 + CT(T_AXF, i + 1) * x11;
      const R l = - R(0.5) * ux_xm;
      const R u = + R(0.5) * ux_xp;
      const R cc = - l - u;
      a -= CT(T_IJDXC, i) * (+ l * wxm + cc * w0 + u * wxp);
    }
    { /* advection_y :123-152 */
      const R uy_ym = CT(T_AYC, i) * y00 + CT(T_AYC, i) * y01;
      const R uy_yp = CT(T_AYC, i) * y10 + CT(T_AYC, i) * y11;
      const R l = - R(0.5) * uy_ym;
      const R u = + R(0.5) * uy_yp;
      const R cc = - l - u;
      a -= CT(T_IJDXC, i) * (+ l * wym + cc * w0 + u * wyp);
    }
    { /* advection_z :154-181 */
      const R uz_zm = CT(T_AZC, i) * wzm + CT(T_AZC, i) * w0;
      const R uz_zp = CT(T_AZC, i) * w0 + CT(T_AZC, i) * wzp;
      const R l = - R(0.5) * uz_zm;
      const R u = + R(0.5) * uz_zp;
      const R cc = - l - u;
      a -= CT(T_IJDXC, i) * (+ l * wzm + cc * w0 + u * wzp);
    }
    { /* diffusion_x0 :183-202 */
      const R v = nu * (+ CT(T_LXC_L, i) * wxm + CT(T_LXC_C, i) * w0 + CT(T_LXC_U, i) * wxp);
      if(ix) gg += v; else a += v;
    }
    { /* diffusion_x1 :204-231; lzx (lzx.c:9-42) at faces i and i+1 */
      const R lzx_xm = R(c.ihz) * (- x00 + x01);
      const R lzx_xp = R(c.ihz) * (- x10 + x11);
      a += CT(T_NUJC, i) * (- CT(T_JHXF, i) * lzx_xm + CT(T_JHXF, i + 1) * lzx_xp);
    }
    { /* diffusion_y0 :233-252 */
      const R v = nu * (+ CT(T_LYC_L, i) * wym + CT(T_LYC_C, i) * w0 + CT(T_LYC_U, i) * wyp);
      if(iy) gg += v; else a += v;
    }
    { /* diffusion_y1 :254-277; lzy (lzy.c:9-42) at j and j+1 */
      const R lzy_ym = R(c.ihz) * (- y00 + y01);
      const R lzy_yp = R(c.ihz) * (- y10 + y11);
      a += CT(T_NUJC, i) * (- CT(T_JHYC, i) * lzy_ym + CT(T_JHYC, i) * lzy_yp);
    }
    { /* diffusion_z :279-299 */
      const R v = nu2 * (+ R(c.lapz_l) * wzm + R(c.lapz_c) * w0 + R(c.lapz_u) * wzp);
      if(iz) gg += v; else a += v;
    }
    /* pressure :301-318 */
    gg -= R(c.ihz) * (- R(Fp_m[of]) + R(Fp_0[of]));
    if(CM > 0) a += R(cvv[2]);
    else if(CM < 0 && A.cz) a += R(A.cz[kb0 + o]);
    R b = R(0.);
    if(sbz) b = sbv ? sbv[2] : sbz[o];
    const R delta = (sbz ? + ca * a + cb * b : + ca * a) + cg * gg;
    if(saz) tck_st(&saz[o], a);
    if(fin) tck_store_halo(nzo - kb0, g, i, j, k, own_value(Fuz_0, of, A.uz, kb0 + o) + (double)delta, false);
    else tck_st(&dzo[o], delta);
  }

  TCK_SECTION_FENCE();
  /* ---------------- T, centres (predict/t.c) ---------------- */
  if(A.f.scalar){
    const R kappa = R(c.kappa);
    R a = R(0.), gg = R(0.);
    const R t0 = R(Ft_0[of]), txm = R(Ft_0[of - 1]), txp = R(Ft_0[of + 1]);
    const R tym = R(Ft_0[of - fsy]), typ = R(Ft_0[of + fsy]);
    const R tzm = R(Ft_m[of]), tzp = R(Ft_p[of]);
    { /* advection_x :88-119 */
      const R ux_xm = CT(T_JHXF, i) * R(Fux_0[of]);
      const R ux_xp = CT(T_JHXF, i + 1) * R(Fux_0[of + 1]);
      const R l = - R(0.5) * ux_xm;
      const R u = + R(0.5) * ux_xp;
      const R cc = - l - u;
      a -= CT(T_IJDXC, i) * (+ l * txm + cc * t0 + u * txp);
    }
    { /* advection_y :121-148 */
      const R uy_ym = CT(T_JHYC, i) * R(Fuy_0[of]);
      const R uy_yp = CT(T_JHYC, i) * R(Fuy_0[of + fsy]);
      const R l = - R(0.5) * uy_ym;
      const R u = + R(0.5) * uy_yp;
      const R cc = - l - u;
      a -= CT(T_IJDXC, i) * (+ l * tym + cc * t0 + u * typ);
    }
    { /* advection_z :150-176 */
      const R uz_zm = CT(T_JHZC, i) * R(Fuz_0[of]);
      const R uz_zp = CT(T_JHZC, i) * R(Fuz_p[of]);
      const R l = - R(0.5) * uz_zm;
      const R u = + R(0.5) * uz_zp;
      const R cc = - l - u;
      a -= CT(T_IJDXC, i) * (+ l * tzm + cc * t0 + u * tzp);
    }
    { /* diffusion_x/y/z :178-239 */
      const R vx = kappa * (+ CT(T_LXC_L, i) * txm + CT(T_LXC_C, i) * t0 + CT(T_LXC_U, i) * txp);
      if(ix) gg += vx; else a += vx;
      const R vy = kappa * (+ CT(T_LYC_L, i) * tym + CT(T_LYC_C, i) * t0 + CT(T_LYC_U, i) * typ);
      if(iy) gg += vy; else a += vy;
      const R vz = kappa * (+ R(c.lapz_l) * tzm + R(c.lapz_c) * t0 + R(c.lapz_u) * tzp);
      if(iz) gg += vz; else a += vz;
    }
    if(CM > 0) a += R(cvv[3]);
    else if(CM < 0 && A.ct) a += R(A.ct[kb0 + o]);
    R b = R(0.);
    if(sbt) b = sbv ? sbv[3] : sbt[o];
    const R delta = (sbt ? + ca * a + cb * b : + ca * a) + cg * gg;
    if(sat) tck_st(&sat[o], a);
    if(fin) tck_store_halo(nto - kb0, g, i, j, k, own_value(Ft_0, of, A.t, kb0 + o) + (double)delta, false);
    else tck_st(&dto[o], delta);
  }
}

/* register budget of the predictor, waves per SIMD: FP64 measures faster at 4 (a few dwords
 * spilled) than at 3, the mixed step's FP32 instantiation slower */
#define TCK_PREDICT_WAVES(R) (sizeof(R) == 8 ? 4 : 3)
template <typename R>
__global__ void __launch_bounds__(TCK_BLOCK) __attribute__((amdgpu_waves_per_eu(TCK_PREDICT_WAVES(R)))) tck_predict(const PredictArgs A){
  __shared__ R smt[T_NPRED * TCK_TW];
  int bx, by, bz;
  tile_of(A.tm, &bx, &by, &bz);
  const int ib = bx * TCK_WAVE;           /* i of LDS column 0: the tile starts at i = ib + 1 */
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int nth = blockDim.x * blockDim.y;
  for(int q = tid; q < T_NPRED * TCK_TW; q += nth){
    const int t = q / TCK_TW;
    const int ii = ib + q % TCK_TW;
    smt[q] = ii < A.c.tp ? R(A.c.tab[(long long)t * A.c.tp + ii]) : R(0.);
  }
  __syncthreads();
  const int i = 1 + ib + threadIdx.x;
  const int j = 1 + by * blockDim.y + threadIdx.y;
  const int k0 = 1 + bz * A.tm.kb;
  const int k1 = min(k0 + A.tm.kb, A.g.nz + 1);
  const double * tf = A.t ? A.t : A.ux;   /* T is read only with the scalar on */
  const long long pl = A.g.plane;
  const R dtv = *A.dt;
  for(int k = k0; k < k1; k++){
    const long long kb0 = (long long)(k - 1) * pl;
    /* field offset within a plane; planes k-1, k, k+1 through three base pointers */
    int of = (i + TC_XOFF) + A.g.px * j;
    asm volatile("" : "+v"(of));
    #define RB(ptr) ((ptr) ? (ptr) + kb0 : nullptr)
    #define RBT(ptr) ((ptr) ? (R *)(ptr) + kb0 : nullptr)
    #define RBC(ptr) ((ptr) ? (const R *)(ptr) + kb0 : nullptr)
    predict_body<R, double, double, TCK_TW, -1>(A, A.ux + kb0, A.ux + kb0 + pl, A.ux + kb0 + 2 * pl,
        A.uy + kb0, A.uy + kb0 + pl, A.uy + kb0 + 2 * pl,
        A.uz + kb0, A.uz + kb0 + pl, A.uz + kb0 + 2 * pl,
        tf + kb0, tf + kb0 + pl, tf + kb0 + 2 * pl,
        A.p + kb0, A.p + kb0 + pl, of, A.g.px,
        RBT(A.dx), RBT(A.dy), RBT(A.dz), RBT(A.dtt),
        RB(A.nx_), RB(A.ny_), RB(A.nz_), RB(A.nt),
        RBC(A.sbx), RBC(A.sby), RBC(A.sbz), RBC(A.sbt), (const R *)nullptr, RBT(A.sax), RBT(A.say), RBT(A.saz), RBT(A.sat),
        i, j, k, kb0, smt, ib, dtv, (const float *)nullptr);
    #undef RB
    #undef RBT
    #undef RBC
  }
}

/* barrier that orders LDS only: global loads issued before it stay in flight and pending
 * global stores are not waited for */
__device__ __forceinline__ void tck_lds_barrier(){
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

/* The same predictor with the advected fields and p staged in LDS.  A block's 256 threads
 * cover a tile of 32 i x 8 j cells and march its kb planes keeping planes k-1, k and k+1 of
 * ux, uy, uz and T (k-1 and k of p) over the tile plus one halo column and row on each side
 * (34 x 10 cells) in a ring of plane buffers, so that every plane is read from memory once
 * per block (1.33 times the tile) instead of once per stencil plane, and the stencils read
 * LDS.  The next plane's loads are issued into registers before the current plane's
 * stencils run.  The planes are staged in the arithmetic type: 49 KiB of LDS in FP64, 25 KiB
 * in the mixed step; three blocks per CU either way (FP64: its LDS; FP32: its registers,
 * held to three waves per SIMD, at which it runs faster than at four with spills).  Outputs
 * and RK sources are addressed as in tck_predict. */
#define TCK_WX 32                          /* tile width in i */
#define TCK_WY 8                           /* tile height in j */
#define TCK_WTW (TCK_WX + 2)               /* window columns (and coefficient-row pitch) */
#define TCK_WPLANE ((TCK_WY + 2) * TCK_WTW)
#define TCK_WPF ((TCK_WPLANE + TCK_BLOCK - 1) / TCK_BLOCK)   /* window elements per thread */


/* SRC: RK sources of the previous substep are read (substeps 1 and 2); CM: closure
 * corrections are added.  Each combination is its own instantiation, so that the registers of
 * what a launch does not read are not held. */
template <typename R, bool SRC, bool CM, int IMP>
__global__ void __launch_bounds__(TCK_BLOCK) __attribute__((amdgpu_waves_per_eu(3))) tck_predict_w(const PredictArgs A){
  __shared__ R smt[T_NPRED * TCK_WTW];
  /* ux uy uz T: a ring of three planes; p: two planes; in the arithmetic type (the stencils
   * read them only through R()).  Each plane has one spare slot past the window, where the
   * threads whose second element lies outside the window put it: the puts are unconditional */
  __shared__ R wf[4][3][TCK_WPLANE + 1];
  __shared__ R wq[2][TCK_WPLANE + 1];
  int bx, by, bz;
  tile_of(A.tm, &bx, &by, &bz);
  const int ib = bx * TCK_WX;             /* i of LDS column 0: the tile starts at i = ib + 1 */
  const int tid = threadIdx.x + TCK_WAVE * threadIdx.y;
  for(int q = tid; q < T_NPRED * TCK_WTW; q += TCK_BLOCK){
    const int t = q / TCK_WTW;
    const int ii = ib + q % TCK_WTW;
    smt[q] = ii < A.c.tp ? R(A.c.tab[(long long)t * A.c.tp + ii]) : R(0.);
  }
  const int li = tid % TCK_WX, lj = tid / TCK_WX;
  const int i = 1 + ib + li;
  const int jw = by * TCK_WY;                                /* j of window row 0 */
  const int j = 1 + jw + lj;
  const int k0 = 1 + bz * A.tm.kb;
  const int k1 = min(k0 + A.tm.kb, A.g.nz + 1);
  const int nrow = min(TCK_WY + 2, A.g.ny + 2 - jw);         /* rows j <= ny + 1 */
  const int ncol = min(TCK_WTW, A.g.px - TC_XOFF - ib);      /* columns inside a row */
  const int nf = A.t ? 4 : 3;
  const long long pl = A.g.plane;
  const long long px = A.g.px;
  /* T's slot reads ux when the scalar is off (loaded, never used) */
  const double * fld[5] = {A.ux, A.uy, A.uz, A.t ? A.t : A.ux, A.p};
  /* this thread's window elements e = tid + q TCK_BLOCK: field offsets within a plane, an
   * in-bounds stand-in (offset 0) for elements outside the field, whose value is replaced by
   * 0 when put; LDS slots, the spare one past the window for e beyond it */
  long long woff[TCK_WPF];
  bool wok[TCK_WPF];
  int wslot[TCK_WPF];
  #pragma unroll
  for(int q = 0; q < TCK_WPF; q++){
    const int e = tid + q * TCK_BLOCK;
    const int r = e / TCK_WTW, cc = e - r * TCK_WTW;
    wok[q] = e < TCK_WPLANE && r < nrow && cc < ncol;
    woff[q] = wok[q] ? (long long)(ib + cc + TC_XOFF) + px * (jw + r) : 0;
    wslot[q] = e < TCK_WPLANE ? e : TCK_WPLANE;
  }
  /* FP64 until the plane is put in LDS: a conversion here would wait for the loads */
  double pre[5][TCK_WPF];
  /* Every global load of the plane loop is unconditional (out-of-range lanes read an
   * in-bounds stand-in and discard it), and the cell's RK sources of the previous substep
   * (substeps 1 and 2) and closure corrections are loaded at the top of each plane, after the
   * staged plane is put in LDS and before the next plane's loads and any store of the plane,
   * and waited for before the first store.  The vector-memory counter counts loads and stores
   * together and the compiler can wait for a load after a store only by draining everything,
   * so this order leaves the next plane's loads in flight through the whole plane; dt is read
   * once. */
  const bool inside = i <= A.g.nx && j <= A.g.ny;
  const long long ocell = (long long)(min(i, A.g.nx) + TC_XOFF) + px * min(j, A.g.ny);
  const R * sbp[4] = {(const R *)A.sbx, (const R *)A.sby, (const R *)A.sbz,
                      (const R *)(A.sbt ? A.sbt : A.sbx)};
  const float * cpp[4] = {A.cx, A.cy, A.cz, A.ct ? A.ct : A.cx};
  const R dtv = *A.dt;
  /* fields at plane kk and p at plane kq into registers */
  #define TCK_FETCH(kk, kq) \
    _Pragma("unroll") for(int f = 0; f < 5; f++) \
      _Pragma("unroll") for(int q = 0; q < TCK_WPF; q++) \
        pre[f][q] = fld[f][woff[q] + pl * (f == 4 ? (kq) : (kk))];
  #define TCK_PUT(slot, pslot) \
    _Pragma("unroll") for(int q = 0; q < TCK_WPF; q++){ \
      _Pragma("unroll") for(int f = 0; f < 4; f++) \
        wf[f][slot][wslot[q]] = (wok[q] && f < nf) ? R(pre[f][q]) : R(0.); \
      wq[pslot][wslot[q]] = wok[q] ? R(pre[4][q]) : R(0.); \
    }
  int sm = 0, s0 = 1, sp = 2, pm = 0;
  TCK_FETCH(k0 - 1, k0 - 1);
  TCK_PUT(sm, pm);
  TCK_FETCH(k0, k0);
  TCK_PUT(s0, 1 - pm);          /* p's plane k0 is written again below */
  TCK_FETCH(k0 + 1, k0);
  const int of = (li + 1) + TCK_WTW * (lj + 1);
  const int ft = nf == 4 ? 3 : 0;
  for(int k = k0; k < k1; k++){
    TCK_PUT(sp, 1 - pm);        /* fields: plane k+1; p: plane k */
    R sbv[4] = {R(0.), R(0.), R(0.), R(0.)};
    float cvv[4] = {0.f, 0.f, 0.f, 0.f};
    if(SRC){
      #pragma unroll
      for(int f = 0; f < 4; f++) sbv[f] = sbp[f][ocell + pl * k];
    }
    if(CM){
      #pragma unroll
      for(int f = 0; f < 4; f++) cvv[f] = cpp[f][ocell + pl * k];
    }
    tck_lds_barrier();
    /* the last plane re-reads plane k + 1 (in bounds) instead of branching around the loads */
    const int kn = k + 1 < k1 ? k + 2 : k + 1;
    TCK_FETCH(kn, kn - 1);
    /* sources and corrections are in registers before the first store */
    if(SRC){ asm volatile("" :: "v"(sbv[0]), "v"(sbv[1]), "v"(sbv[2]), "v"(sbv[3])); }
    if(CM){ asm volatile("" :: "v"(cvv[0]), "v"(cvv[1]), "v"(cvv[2]), "v"(cvv[3])); }

    const long long kb0 = (long long)(k - 1) * pl;
    #define RB(ptr) ((ptr) ? (ptr) + kb0 : nullptr)
    #define RBT(ptr) ((ptr) ? (R *)(ptr) + kb0 : nullptr)
    #define RBC(ptr) ((ptr) ? (const R *)(ptr) + kb0 : nullptr)
    if(inside)
    predict_body<R, R, R, TCK_WTW, CM ? 1 : 0, IMP>(A, wf[0][sm], wf[0][s0], wf[0][sp], wf[1][sm], wf[1][s0], wf[1][sp],
        wf[2][sm], wf[2][s0], wf[2][sp], wf[ft][sm], wf[ft][s0], wf[ft][sp],
        wq[pm], wq[1 - pm], of, TCK_WTW,
        RBT(A.dx), RBT(A.dy), RBT(A.dz), RBT(A.dtt),
        RB(A.nx_), RB(A.ny_), RB(A.nz_), RB(A.nt),
        SRC ? RBC(A.sbx) : nullptr, SRC ? RBC(A.sby) : nullptr, SRC ? RBC(A.sbz) : nullptr,
        SRC ? RBC(A.sbt) : nullptr, (const R *)sbv,
        RBT(A.sax), RBT(A.say), RBT(A.saz), RBT(A.sat), i, j, k, kb0, smt, ib, dtv, (const float *)cvv);
    #undef RB
    #undef RBT
    #undef RBC
    tck_lds_barrier();
    const int tmp = sm; sm = s0; s0 = sp; sp = tmp;
    pm = 1 - pm;
  }
  #undef TCK_FETCH
  #undef TCK_PUT
}

/* ------------------------------------------------------------------------------------- */

struct CorrectArgs {
  tc_geom_t g;
  tc_coef_t c;
  tc_flags_t f;
  double * ux; double * uy; double * uz; double * p;
  const void * psi;             /* the potential: double, or float in the mixed step */
  const double * dt;
  int rkstep;
  int mode;                     /* bit 0: velocity correction, bit 1: pressure update */
  unsigned long long * dtmin;   /* non-null: CFL minimum of the corrected field */
  TileMap tm;
};

/* Per-column constants of the correction: everything in correct/*.c and update_pressure.c
 * that depends on i and the substep only (the reference's left-to-right evaluation of
 * prefactor / HXXF(i) * (...) and 2 f / J * (...) makes these factors exact subexpressions),
 * evaluated once per thread before its plane loop. */
struct CorrectCol {
  double fx, fy, fz;        /* gamma dt / HXXF(i), gamma dt / HYXC(i), gamma dt / hz */
  double hx, hy;            /* HXXF(i), HYXC(i) (CFL bounds) */
  double qf;                /* 2 f / JDXC(i), f = gamma dt nu / 2 */
  double xl, xu, yl, zl;   /* implicit-direction Laplacian of psi (centre: - l - u) */
};

/* gamma dt / hz (uniform) */
__device__ __forceinline__ double correct_col_fz(const CorrectArgs & A){
  const int s = A.rkstep;
  const double gamma = c_rka[s] + c_rkb[s];
  const double prefactor = gamma * *A.dt;
  return prefactor / A.c.hz;
}

__device__ __forceinline__ CorrectCol correct_col(const CorrectArgs & A, int i){
  const tc_coef_t & c = A.c;
  CorrectCol C;
  const double dt = *A.dt;
  const int s = A.rkstep;
  const double gamma = c_rka[s] + c_rkb[s];
  const double prefactor = gamma * dt;                /* correct/main.c:22-23 */
  C.hx = c.hxxf[i];
  C.hy = c.hyxc[i];
  C.fx = prefactor / C.hx;
  C.fy = prefactor / C.hy;
  C.fz = prefactor / c.hz;
  const double pf = 0.5 * gamma * dt * c.nu;          /* update_pressure.c:155-156 */
  C.qf = 2. * pf / c.jdxc[i];
  C.xl = c.plx_l[i]; C.xu = c.plx_u[i];
  C.yl = c.ply[i];
  C.zl = c.plz[i];
  return C;
}

/* one cell of the correction.  Every load of the cell is issued before its first store (the
 * vector-memory counter counts loads and stores together, so a load issued after a store
 * could be waited for only by draining both).  The fields are read and written through
 * buffer resources of plane k and psi through one of planes k-1 .. k+1 based one row and one
 * cell before the plane, so that every address is one per-lane offset (ob: the cell's byte
 * offset in a plane of doubles, obr: in a plane of R) plus a uniform one.  psi of the cell
 * at planes k and k - 1 comes in from the previous plane (ps0, pszm); the one at k + 1 goes
 * out for the next (pszp_out).  CFL: also the cell's bound of decide_dt_adv on the
 * corrected velocity. */
template <typename R, bool CFL>
__device__ __forceinline__ double correct_cell(const CorrectArgs & A, const CorrectCol & C,
    const int i, const int j, const int k, const unsigned ob, const unsigned obr,
    const double ps0, const double pszm, double & pszp_out){
  const tc_geom_t & g = A.g;
  const long long pl = g.plane;
  const int sy = g.px;
  const long long kk = (long long)k * pl;
  const unsigned fb = (unsigned)pl * 8u;
  const __amdgpu_buffer_rsrc_t rux = tck_rsrc(A.ux + kk, fb), ruy = tck_rsrc(A.uy + kk, fb),
      ruz = tck_rsrc(A.uz + kk, fb), rp = tck_rsrc(A.p + kk, fb);
  const unsigned sr = (unsigned)sizeof(R);
  const __amdgpu_buffer_rsrc_t rps = tck_rsrc((const R *)A.psi + kk - pl - sy - 1,
      (unsigned)(3 * pl + 2 * sy + 2) * sr);
  const unsigned c0 = (unsigned)(pl + sy + 1) * sr;        /* the cell itself */
  double lmin = 1.;
  const double small = 1.e-8;
  const bool vel = A.mode & 1, prs = A.mode & 2;
  const bool ix = prs && A.f.implicit[0], iy = prs && A.f.implicit[1], iz = prs && A.f.implicit[2];
  /* all loads unconditional (a launch that skips the velocity or the pressure part reads a
   * field it does not use): no branch between them, no register cleared in between */
  const double psxm = double(tck_bld<R>(rps, obr, c0 - sr));
  const double psym = double(tck_bld<R>(rps, obr, c0 - (unsigned)sy * sr));
  const double psxp = double(tck_bld<R>(rps, obr, c0 + sr));
  const double psyp = double(tck_bld<R>(rps, obr, c0 + (unsigned)sy * sr));
  const double pszp = double(tck_bld<R>(rps, obr, c0 + (unsigned)pl * sr));
  pszp_out = pszp;
  const double u0 = tck_bld<double>(rux, ob, 0), v0 = tck_bld<double>(ruy, ob, 0);
  const double w0 = tck_bld<double>(ruz, ob, 0), p0 = tck_bld<double>(rp, ob, 0);
  /* halo images (src/halo.c:18-142: j = 0 from ny, j = ny+1 from 1, then the z planes over the
   * full (x, y) extent, so corners are images of images): row nj of plane k, and plane nk */
  const int nj = (j == 1) ? g.ny + 1 : (j == g.ny) ? 0 : -1;
  const int nk = (k == 1) ? g.nz + 1 : (k == g.nz) ? 0 : -1;
  const unsigned obj = ob + (unsigned)((nj - j) * sy) * 8u;
  const long long kn = (long long)nk * pl;
  double vx = 0., vy = 0., vz = 0., vp = 0.;
  if(vel){
    if(i >= 2){ /* correct/ux.c:25-52 */
      vx = u0 - C.fx * (- psxm + ps0);
      tck_bst(rux, ob, vx);
      if(CFL) lmin = fmin(lmin, C.hx / (fabs(vx) + small));
    }
    { /* correct/uy.c:25-52 */
      vy = v0 - C.fy * (- psym + ps0);
      tck_bst(ruy, ob, vy);
      if(CFL) lmin = fmin(lmin, C.hy / (fabs(vy) + small));
    }
    { /* correct/uz.c:15-44 */
      vz = w0 - C.fz * (- pszm + ps0);
      tck_bst(ruz, ob, vz);
      if(CFL) lmin = fmin(lmin, A.c.hz / (fabs(vz) + small));
    }
  }
  if(prs){ /* update_pressure.c:146-173 */
    vp = p0 + ps0;
    if(ix) vp -= C.qf * (+ C.xl * psxm + (- C.xl - C.xu) * ps0 + C.xu * psxp);
    if(iy) vp -= C.qf * (+ C.yl * psym + (- C.yl - C.yl) * ps0 + C.yl * psyp);
    if(iz) vp -= C.qf * (+ C.zl * pszm + (- C.zl - C.zl) * ps0 + C.zl * pszp);
    tck_bst(rp, ob, vp);
  }
  if(nj >= 0){
    if(vel){
      if(i >= 2) tck_bsti(rux, obj, vx);
      tck_bsti(ruy, obj, vy);
      tck_bsti(ruz, obj, vz);
    }
    if(prs) tck_bsti(rp, obj, vp);
  }
  if(nk >= 0){
    const __amdgpu_buffer_rsrc_t nux = tck_rsrc(A.ux + kn, fb), nuy = tck_rsrc(A.uy + kn, fb),
        nuz = tck_rsrc(A.uz + kn, fb), np = tck_rsrc(A.p + kn, fb);
    if(vel){
      if(i >= 2){ tck_bsti(nux, ob, vx); if(nj >= 0) tck_bsti(nux, obj, vx); }
      tck_bsti(nuy, ob, vy); if(nj >= 0) tck_bsti(nuy, obj, vy);
      tck_bsti(nuz, ob, vz); if(nj >= 0) tck_bsti(nuz, obj, vz);
    }
    if(prs){ tck_bsti(np, ob, vp); if(nj >= 0) tck_bsti(np, obj, vp); }
  }
  /* the x wall copy of p (src/fluid/boundary/p.c): i = 0 from 1, i = nx + 1 from nx, with its
   * own images */
  if(prs && (i == 1 || i == g.nx)){
    const unsigned dx = i == 1 ? 0u - 8u : 8u;
    tck_bst(rp, ob + dx, vp);
    if(nj >= 0) tck_bsti(rp, obj + dx, vp);
    if(nk >= 0){
      const __amdgpu_buffer_rsrc_t np = tck_rsrc(A.p + kn, fb);
      tck_bsti(np, ob + dx, vp);
      if(nj >= 0) tck_bsti(np, obj + dx, vp);
    }
  }
  return lmin;
}

/* block minimum of v folded into one of the TCK_DTSLOTS accumulators */
__device__ __forceinline__ void dt_block_min(double v, unsigned long long * dtmin){
  __shared__ double red[TCK_BLOCK / TCK_WAVE];
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  double m = tck_wave_min(v);
  if((tid % TCK_WAVE) == 0) red[tid / TCK_WAVE] = m;
  __syncthreads();
  if(tid == 0){
    const int nw = (blockDim.x * blockDim.y + TCK_WAVE - 1) / TCK_WAVE;
    for(int w = 1; w < nw; w++) m = fmin(m, red[w]);
    const long long bid = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    atomicMin(dtmin + (bid % TCK_DTSLOTS) * TCK_DTSTRIDE, tck_bits(m));
  }
}

/* the per-column factors of a block's 64 columns, in LDS: read again for every plane rather
 * than held in registers, which leaves the kernel at 8 waves per SIMD */
#define TCK_CCOLS 9
template <typename R, bool CFL>
__global__ void __launch_bounds__(TCK_BLOCK) tck_correct(const CorrectArgs A){
  __shared__ double colv[TCK_CCOLS][TCK_WAVE];
  int bx, by, bz;
  tile_of(A.tm, &bx, &by, &bz);
  const int i = 1 + bx * TCK_WAVE + threadIdx.x;
  const int j = 1 + by * blockDim.y + threadIdx.y;
  const int k0 = 1 + bz * A.tm.kb;
  const int k1 = min(k0 + A.tm.kb, A.g.nz + 1);
  if(threadIdx.y == 0){
    const CorrectCol C = correct_col(A, min(i, A.g.nx));
    const double v[TCK_CCOLS] = {C.fx, C.fy, C.qf, C.xl, C.xu, C.yl, C.zl, C.hx, C.hy};
    #pragma unroll
    for(int q = 0; q < TCK_CCOLS; q++) colv[q][threadIdx.x] = v[q];
  }
  __syncthreads();
  double lmin = 1.;
  if(i <= A.g.nx && j <= A.g.ny){
    CorrectCol C;
    C.fz = correct_col_fz(A);
    /* the cell's offset within a plane (i <= nx + 1, j <= ny + 1: far below 2^32 bytes) */
    const unsigned oc = (unsigned)((i + TC_XOFF) + A.g.px * j);
    const unsigned ob = oc * 8u, obr = oc * (unsigned)sizeof(R);
    /* psi of the cell at planes k - 1 and k, carried from plane to plane: each plane of psi
     * is read from memory once per column of cells */
    const R * psc = (const R *)A.psi + oc;
    double pszm = double(psc[(long long)(k0 - 1) * A.g.plane]), ps0 = double(psc[(long long)k0 * A.g.plane]);
    for(int k = k0; k < k1; k++){
      /* keeps the column factors from being hoisted out of the loop into registers */
      asm volatile("" ::: "memory");
      const int x = threadIdx.x;
      C.fx = colv[0][x]; C.fy = colv[1][x]; C.qf = colv[2][x]; C.xl = colv[3][x]; C.xu = colv[4][x];
      C.yl = colv[5][x]; C.zl = colv[6][x]; C.hx = colv[7][x]; C.hy = colv[8][x];
      double pszp;
      lmin = fmin(lmin, correct_cell<R, CFL>(A, C, i, j, k, ob, obr, ps0, pszm, pszp));
      pszm = ps0;
      ps0 = pszp;
    }
  }
  /* CFL bound of decide_dt_adv (decide_dt.c:32-80) on the corrected velocity */
  if(CFL) dt_block_min(lmin, A.dtmin);
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
  dt_block_min(lmin, A.dtmin);
}

/* decide_dt.c:125-160: dt = coef_adv * min(1, CFL) then min with the explicit diffusive
 * limits (precomputed on the host, decide_dt.c:89-115); re-arms the CFL accumulators.
 * One wave: lane t owns accumulator t. */
__global__ void tck_dt_finalise(unsigned long long * dtmin, double * dt, double coef_adv,
    double dif_limit, double dt_fixed){
  const int t = threadIdx.x;
  double v = 1.;
  if(t < TCK_DTSLOTS){
    v = __longlong_as_double((long long)dtmin[t * TCK_DTSTRIDE]);
    dtmin[t * TCK_DTSTRIDE] = tck_bits(1.);
  }
  v = tck_wave_min(v);
  if(t == 0){
    if(dt_fixed > 0.){
      *dt = dt_fixed;
    }else{
      v *= coef_adv;
      *dt = fmin(v, dif_limit);
    }
  }
}

__global__ void tck_dt_arm(unsigned long long * dtmin){
  if(threadIdx.x < TCK_DTSLOTS) dtmin[threadIdx.x * TCK_DTSTRIDE] = tck_bits(1.);
}

/* ------------------------------------------------------------------------------------- */
extern "C" {

static TileMap cell_tiles(const tc_geom_t & g, int kb, int ty){
  TileMap T;
  T.nbx = (g.nx + TCK_WAVE - 1) / TCK_WAVE;
  T.nby = (g.ny + ty - 1) / ty;
  T.kb = kb < 1 ? 1 : kb;
  T.nbz = (g.nz + T.kb - 1) / T.kb;
  T.total = (long long)T.nbx * T.nby * T.nbz;
  return T;
}

static dim3 cell_grid(const tc_geom_t & g, dim3 * blk){
  *blk = dim3(64, 4, 1);
  return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
}

static int check_shape(int kb, int ty){
  return (kb >= 1 && (ty == 1 || ty == 2 || ty == 4)) ? 0 : (int)hipErrorInvalidValue;
}

int tck_launch_predict(const tc_geom_t * g, const tc_coef_t * c, const tc_flags_t * f,
    const double * const cur[4], const double * p, double * const nxt[4], void * const dlt[4],
    const void * const srcb[4], void * const srca[4], const float * const corr[4],
    const double * dt, int rkstep, int final_here, int mixed, int kb, int ty, int window, hipStream_t stream){
  if(check_shape(kb, ty)) return (int)hipErrorInvalidValue;
  if(window && ty != 4) window = 0;       /* the LDS-window variant goes with 4-row shapes */
  PredictArgs A;
  A.g = *g; A.c = *c; A.f = *f;
  A.ux = cur[0]; A.uy = cur[1]; A.uz = cur[2]; A.t = cur[3]; A.p = p;
  A.nx_ = nxt[0]; A.ny_ = nxt[1]; A.nz_ = nxt[2]; A.nt = nxt[3];
  A.dx = dlt[0]; A.dy = dlt[1]; A.dz = dlt[2]; A.dtt = dlt[3];
  A.sbx = srcb[0]; A.sby = srcb[1]; A.sbz = srcb[2]; A.sbt = srcb[3];
  A.sax = srca[0]; A.say = srca[1]; A.saz = srca[2]; A.sat = srca[3];
  A.cx = corr ? corr[0] : nullptr; A.cy = corr ? corr[1] : nullptr;
  A.cz = corr ? corr[2] : nullptr; A.ct = corr ? corr[3] : nullptr;
  A.dt = dt; A.rkstep = rkstep; A.final_here = final_here;
  A.tm = cell_tiles(*g, kb, ty);
  if(window){
    /* 32 x 8 tiles of 256 threads */
    A.tm.nbx = (g->nx + TCK_WX - 1) / TCK_WX;
    A.tm.nby = (g->ny + TCK_WY - 1) / TCK_WY;
    A.tm.total = (long long)A.tm.nbx * A.tm.nby * A.tm.nbz;
  }
  const dim3 grd((unsigned)A.tm.total), blk(TCK_WAVE, window ? TCK_BLOCK / TCK_WAVE : ty);
  if(window){
    const bool src = A.sbx || A.sby || A.sbz || A.sbt;
    const bool cm = A.cx || A.cy || A.cz || A.ct;
    /* the all-implicit step (the benchmark's) has its own instantiations: no flag selects,
     * no final-store path, fewer scalar registers */
    const bool all3 = !final_here && f->implicit[0] && f->implicit[1] && f->implicit[2];
    #define TCK_PW(RT, IM) \
      if(src){ if(cm) hipLaunchKernelGGL((tck_predict_w<RT, true, true, IM>), grd, blk, 0, stream, A); \
               else hipLaunchKernelGGL((tck_predict_w<RT, true, false, IM>), grd, blk, 0, stream, A); } \
      else   { if(cm) hipLaunchKernelGGL((tck_predict_w<RT, false, true, IM>), grd, blk, 0, stream, A); \
               else hipLaunchKernelGGL((tck_predict_w<RT, false, false, IM>), grd, blk, 0, stream, A); }
    if(mixed){ if(all3){ TCK_PW(float, 7) } else { TCK_PW(float, -1) } }
    else { if(all3){ TCK_PW(double, 7) } else { TCK_PW(double, -1) } }
    #undef TCK_PW
  }else{
    if(mixed) hipLaunchKernelGGL(tck_predict<float>, grd, blk, 0, stream, A);
    else hipLaunchKernelGGL(tck_predict<double>, grd, blk, 0, stream, A);
  }
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_correct(const tc_geom_t * g, const tc_coef_t * c, const tc_flags_t * f,
    double * ux, double * uy, double * uz, double * p, const void * psi, const double * dt,
    int rkstep, int mode, unsigned long long * dtmin, int mixed, int kb, int ty, hipStream_t stream){
  if(check_shape(kb, ty)) return (int)hipErrorInvalidValue;
  /* buffer resources span three planes of psi: their size must fit the 31-bit record count */
  if((3 * g->plane + 2 * (long long)g->px + 2) * 8 >= (1LL << 31)) return (int)hipErrorInvalidValue;
  CorrectArgs A;
  A.mode = mode;
  A.g = *g; A.c = *c; A.f = *f;
  A.ux = ux; A.uy = uy; A.uz = uz; A.p = p; A.psi = psi;
  A.dt = dt; A.rkstep = rkstep; A.dtmin = dtmin;
  A.tm = cell_tiles(*g, kb, ty);
  const dim3 grid((unsigned)A.tm.total), block(TCK_WAVE, ty);
  if(mixed){
    if(dtmin) hipLaunchKernelGGL((tck_correct<float, true>), grid, block, 0, stream, A);
    else hipLaunchKernelGGL((tck_correct<float, false>), grid, block, 0, stream, A);
  } else {
    if(dtmin) hipLaunchKernelGGL((tck_correct<double, true>), grid, block, 0, stream, A);
    else hipLaunchKernelGGL((tck_correct<double, false>), grid, block, 0, stream, A);
  }
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_dt(const tc_geom_t * g, const tc_coef_t * c, const double * ux, const double * uy,
    const double * uz, unsigned long long * dtmin, hipStream_t stream){
  DtArgs A;
  A.g = *g; A.c = *c; A.ux = ux; A.uy = uy; A.uz = uz; A.dtmin = dtmin;
  hipLaunchKernelGGL(tck_dt_arm, dim3(1), dim3(TCK_WAVE), 0, stream, dtmin);
  TCK_CHECK_LAUNCH();
  dim3 blk;
  dim3 grd = cell_grid(*g, &blk);
  hipLaunchKernelGGL(tck_dt_reduce, grd, blk, 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_dt_finalise(unsigned long long * dtmin, double * dt, double coef_adv,
    double dif_limit, double dt_fixed, hipStream_t stream){
  hipLaunchKernelGGL(tck_dt_finalise, dim3(1), dim3(TCK_WAVE), 0, stream, dtmin, dt, coef_adv, dif_limit, dt_fixed);
  TCK_CHECK_LAUNCH();
  return 0;
}

}
