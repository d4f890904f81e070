/*
 * ORACLE (test infrastructure only) -- body of the CPU restatement of the
 * reference "Hand" DWF Wilson hopping kernel, instantiated once per real type
 * by hand.c (OR_T = double | float, OR_SFX = f64 | f32).
 *
 * Follows, operation for operation:
 *   leg order and per-leg projector/reconstruct pairs  WilsonKernelsHandCpu.h:426-446
 *   HAND_STENCIL_LEG (table read, load, project, permute, mult, recon)
 *                                                     WilsonKernelsHandCpu.h:307-317
 *   projectors *_PROJ                                 WilsonKernelsHandCpu.h:94-159
 *   MULT_2SPIN (column 0, then 1, then 2)             WilsonKernelsHandCpu.h:62-90
 *   reconstructions *_RECON / *_RECON_ACCUM           WilsonKernelsHandCpu.h:165-303
 *   HAND_RESULT (12 stores, out fully overwritten)    WilsonKernelsHandCpu.h:319-334
 *   complex mult  re=ar*br-ai*bi, im=ar*bi+ai*br      arch/generic/Simd_generic.h:153-155
 *   timesI (-im,re), timesMinusI (im,-re)             arch/generic/Simd_generic.h:238-240,259-261
 *   lane permute out[l] = in[l ^ (Nsimd>>(ptype+1))]  WilsonKernelsHandGpu.h:36-37,
 *                                                     arch/gpu/Simd_gpu_vec.h:441-451
 * Must be compiled with -ffp-contract=off so every product is rounded before
 * the add/sub that consumes it, as in the reference's generic backend.
 */

typedef struct { OR_T re, im; } OR_CAT(cplx_, OR_SFX);
#define CT OR_CAT(cplx_, OR_SFX)

static inline CT OR_CAT(cadd_, OR_SFX)(CT a, CT b) { CT r; r.re = a.re + b.re; r.im = a.im + b.im; return r; }
static inline CT OR_CAT(csub_, OR_SFX)(CT a, CT b) { CT r; r.re = a.re - b.re; r.im = a.im - b.im; return r; }
static inline CT OR_CAT(cmul_, OR_SFX)(CT a, CT b) {
  CT r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}
static inline CT OR_CAT(ctI_, OR_SFX)(CT a) { CT r; r.re = -a.im; r.im = a.re; return r; }
static inline CT OR_CAT(ctmI_, OR_SFX)(CT a) { CT r; r.re = a.im; r.im = -a.re; return r; }

#define CADD OR_CAT(cadd_, OR_SFX)
#define CSUB OR_CAT(csub_, OR_SFX)
#define CMUL OR_CAT(cmul_, OR_SFX)
#define CTI  OR_CAT(ctI_, OR_SFX)
#define CTMI OR_CAT(ctmI_, OR_SFX)

/* Spin projection 4 -> 2 (upper two rows of (1 +/- gamma_mu) psi in the
 * chiral basis).  WilsonKernelsHandCpu.h:94-159. */
static void OR_CAT(proj_, OR_SFX)(int kind, const CT p[4][3], CT h[2][3])
{
  for (int c = 0; c < 3; c++) {
    switch (kind) {
    case OR_XP: h[0][c] = CADD(p[0][c], CTI(p[3][c])); h[1][c] = CADD(p[1][c], CTI(p[2][c])); break;
    case OR_YP: h[0][c] = CSUB(p[0][c], p[3][c]);      h[1][c] = CADD(p[1][c], p[2][c]);      break;
    case OR_ZP: h[0][c] = CADD(p[0][c], CTI(p[2][c])); h[1][c] = CSUB(p[1][c], CTI(p[3][c])); break;
    case OR_TP: h[0][c] = CADD(p[0][c], p[2][c]);      h[1][c] = CADD(p[1][c], p[3][c]);      break;
    case OR_XM: h[0][c] = CSUB(p[0][c], CTI(p[3][c])); h[1][c] = CSUB(p[1][c], CTI(p[2][c])); break;
    case OR_YM: h[0][c] = CADD(p[0][c], p[3][c]);      h[1][c] = CSUB(p[1][c], p[2][c]);      break;
    case OR_ZM: h[0][c] = CSUB(p[0][c], CTI(p[2][c])); h[1][c] = CADD(p[1][c], CTI(p[3][c])); break;
    default:    h[0][c] = CSUB(p[0][c], p[2][c]);      h[1][c] = CSUB(p[1][c], p[3][c]);      break; /* TM */
    }
  }
}

/* Spin reconstruction 2 -> 4.  first != 0 selects the assigning form
 * (XP_RECON / XM_RECON), otherwise *_RECON_ACCUM.
 * WilsonKernelsHandCpu.h:165-303. */
static void OR_CAT(recon_, OR_SFX)(int kind, int first, const CT u[2][3], CT r[4][3])
{
  for (int c = 0; c < 3; c++) {
    if (first) {
      r[0][c] = u[0][c];
      r[1][c] = u[1][c];
      if (kind == OR_XP) { r[2][c] = CTMI(u[1][c]); r[3][c] = CTMI(u[0][c]); }
      else               { r[2][c] = CTI(u[1][c]);  r[3][c] = CTI(u[0][c]);  } /* XM */
      continue;
    }
    r[0][c] = CADD(r[0][c], u[0][c]);
    r[1][c] = CADD(r[1][c], u[1][c]);
    switch (kind) {
    case OR_XP: r[2][c] = CSUB(r[2][c], CTI(u[1][c])); r[3][c] = CSUB(r[3][c], CTI(u[0][c])); break;
    case OR_XM: r[2][c] = CADD(r[2][c], CTI(u[1][c])); r[3][c] = CADD(r[3][c], CTI(u[0][c])); break;
    case OR_YP: r[2][c] = CADD(r[2][c], u[1][c]);      r[3][c] = CSUB(r[3][c], u[0][c]);      break;
    case OR_YM: r[2][c] = CSUB(r[2][c], u[1][c]);      r[3][c] = CADD(r[3][c], u[0][c]);      break;
    case OR_ZP: r[2][c] = CSUB(r[2][c], CTI(u[0][c])); r[3][c] = CADD(r[3][c], CTI(u[1][c])); break;
    case OR_ZM: r[2][c] = CADD(r[2][c], CTI(u[0][c])); r[3][c] = CSUB(r[3][c], CTI(u[1][c])); break;
    case OR_TP: r[2][c] = CADD(r[2][c], u[0][c]);      r[3][c] = CADD(r[3][c], u[1][c]);      break;
    default:    r[2][c] = CSUB(r[2][c], u[0][c]);      r[3][c] = CSUB(r[3][c], u[1][c]);      break; /* TM */
    }
  }
}

/*
 * One application of the hopping term on every 5D outer site and lane.
 *   U   [nsite][8][3][3][nsimd][2]   doubled gauge field (reference host layout)
 *   in  [nsite*Ls][4][3][nsimd][2]   input spinor
 *   out [nsite*Ls][4][3][nsimd][2]   output spinor (fully overwritten)
 *   nbr [nsite*Ls*8], prm [nsite*Ls*8]  stencil table, entry ss*8+dir
 * dag = 0 is the reference kernel; dag = 1 swaps every projector/reconstruct
 * pair (forward legs *P_*, backward legs *M_*), which is gamma5 Dhop gamma5.
 * nthreads > 1 splits the 4D outer-site loop over OpenMP threads, as the
 * reference's own OpenMP build does (WilsonKernelsHandCpu.h:416-426).
 */
void OR_CAT(oracle_dhop_hand_, OR_SFX)(int nsimd, int dag, const OR_T *U, OR_T *out, const OR_T *in,
                                       const uint64_t *nbr, const uint8_t *prm,
                                       uint64_t nsite, uint64_t Ls, int nthreads)
{
  /* reference leg table: direction, projector, permute type, reconstruct */
  static const int dir_of_leg[8]   = {0, 1, 2, 3, 4, 5, 6, 7};          /* Xp..Tm */
  static const int ptype_of_leg[8] = {3, 2, 1, 0, 3, 2, 1, 0};
  static const int proj_nodag[8]   = {OR_XM, OR_YM, OR_ZM, OR_TM, OR_XP, OR_YP, OR_ZP, OR_TP};
  static const int proj_dag[8]     = {OR_XP, OR_YP, OR_ZP, OR_TP, OR_XM, OR_YM, OR_ZM, OR_TM};
  const int *proj = dag ? proj_dag : proj_nodag;
  const uint64_t sp_stride = (uint64_t)24 * nsimd;       /* reals per 5D outer site */
  const uint64_t u_stride  = (uint64_t)8 * 18 * nsimd;   /* reals per 4D outer site */
  (void)nthreads;

#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int64_t ssite = 0; ssite < (int64_t)nsite; ssite++) {
    const uint64_t sU = (uint64_t)ssite;
    for (uint64_t s = 0; s < Ls; s++) {
      const uint64_t ss = sU * Ls + s;
      for (int lane = 0; lane < nsimd; lane++) {
        CT res[4][3];
        for (int leg = 0; leg < 8; leg++) {
          const int dir = dir_of_leg[leg];
          const uint64_t offset = nbr[ss * 8 + dir];
          const int perm = prm[ss * 8 + dir];
          const int mask = nsimd >> (ptype_of_leg[leg] + 1);
          const int plane = perm ? (lane ^ mask) : lane;
          CT p[4][3], h[2][3], uh[2][3];
          const OR_T *src = in + offset * sp_stride;
          for (int sp = 0; sp < 4; sp++)
            for (int c = 0; c < 3; c++) {
              const OR_T *q = src + ((uint64_t)(sp * 3 + c) * nsimd + plane) * 2;
              p[sp][c].re = q[0];
              p[sp][c].im = q[1];
            }
          OR_CAT(proj_, OR_SFX)(proj[leg], p, h);
          /* MULT_2SPIN: UChi[sp][i] = ((U_i0 chi_0) + U_i1 chi_1) + U_i2 chi_2 */
          const OR_T *ub = U + sU * u_stride + (uint64_t)dir * 18 * nsimd;
          CT u[3][3];
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
              const OR_T *q = ub + ((uint64_t)(i * 3 + j) * nsimd + lane) * 2;
              u[i][j].re = q[0];
              u[i][j].im = q[1];
            }
          for (int sp = 0; sp < 2; sp++)
            for (int i = 0; i < 3; i++) {
              CT t = CMUL(u[i][0], h[sp][0]);
              t = CADD(t, CMUL(u[i][1], h[sp][1]));
              t = CADD(t, CMUL(u[i][2], h[sp][2]));
              uh[sp][i] = t;
            }
          OR_CAT(recon_, OR_SFX)(proj[leg], leg == 0, uh, res);
        }
        OR_T *dst = out + ss * sp_stride;
        for (int sp = 0; sp < 4; sp++)
          for (int c = 0; c < 3; c++) {
            OR_T *q = dst + ((uint64_t)(sp * 3 + c) * nsimd + lane) * 2;
            q[0] = res[sp][c].re;
            q[1] = res[sp][c].im;
          }
      }
    }
  }
}

#undef CT
#undef CADD
#undef CSUB
#undef CMUL
#undef CTI
#undef CTMI
