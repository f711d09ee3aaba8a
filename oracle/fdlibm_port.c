/*
 * ORACLE — test infrastructure only. Nothing in the product links this file; tests, smoke() and the
 * CPU-baseline leg of bench.py load it through liboracle.so as the checker.
 *
 * A plain-C restatement of the netlib fdlibm 5.3 routines the flooding step uses: sin, cos (with their
 * kernels and the small/medium argument reduction), pow, scalbn and copysign. The algorithm and its
 * constants are fdlibm's (Sun Microsystems, freely distributable); the reference vendors the C sources in
 * third_party/fdlibm/src and restates them as host+device code in
 * engine/include/sinksim/kernel/fdlibm_portable.hpp:20-435.
 *
 * Word access goes through memcpy so that the code is free of the aliasing tricks of the original.
 * Arguments of sin/cos above 2^19*pi/2 would need fdlibm's large-argument reduction (k_rem_pio2.c); like
 * the reference (fdlibm_portable.hpp:224,239) this falls back to the platform function there. The step
 * clamps its angles to 1.4 rad, far inside the covered range.
 *
 * Parity pin: tests/test_oracle_fdlibm.py compares every function bit for bit with the netlib build of
 * the reference's vendored sources (oracle/_ref/libfdlibm_ref.so) on the SplitMix64 sample recipe of
 * tests/test_portable_math.cpp:40-105, and with the committed vector fixture tests/golden/fdlibm_vectors.npz.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

static int32_t hi_of(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return (int32_t)(uint32_t)(b >> 32);
}
static uint32_t lo_of(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return (uint32_t)b;
}
static double words(int32_t h, uint32_t l) {
  uint64_t b = ((uint64_t)(uint32_t)h << 32) | (uint64_t)l;
  double x;
  memcpy(&x, &b, 8);
  return x;
}
static double set_hi(double x, int32_t h) { return words(h, lo_of(x)); }
static double set_lo(double x, uint32_t l) { return words(hi_of(x), l); }

/* s_copysign.c */
double orc_fdl_copysign(double x, double y) {
  uint32_t h = ((uint32_t)hi_of(x) & 0x7fffffffu) | ((uint32_t)hi_of(y) & 0x80000000u);
  return set_hi(x, (int32_t)h);
}

/* s_scalbn.c */
double orc_fdl_scalbn(double x, int n) {
  const double two54 = 1.80143985094819840000e+16, twom54 = 5.55111512312578270212e-17;
  const double huge = 1.0e+300, tiny = 1.0e-300;
  int32_t hx = hi_of(x);
  uint32_t lx = lo_of(x);
  int32_t k = (hx & 0x7ff00000) >> 20;
  if (k == 0) {
    if ((lx | (uint32_t)(hx & 0x7fffffff)) == 0) return x;
    x *= two54;
    hx = hi_of(x);
    k = ((hx & 0x7ff00000) >> 20) - 54;
    if (n < -50000) return tiny * x;
  }
  if (k == 0x7ff) return x + x;
  k = k + n;
  if (k > 0x7fe) return huge * orc_fdl_copysign(huge, x);
  if (k > 0) return set_hi(x, (int32_t)(((uint32_t)hx & 0x800fffffu) | ((uint32_t)k << 20)));
  if (k <= -54) {
    if (n > 50000) return huge * orc_fdl_copysign(huge, x);
    return tiny * orc_fdl_copysign(tiny, x);
  }
  k += 54;
  x = set_hi(x, (int32_t)(((uint32_t)hx & 0x800fffffu) | ((uint32_t)k << 20)));
  return x * twom54;
}

/* k_sin.c: sin on [-pi/4, pi/4]; y is the tail of x, iy = 0 when y is zero */
static double k_sin(double x, double y, int iy) {
  const double half = 5.00000000000000000000e-01;
  const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03, S3 = -1.98412698298579493134e-04;
  const double S4 = 2.75573137070700676789e-06, S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
  int32_t ix = hi_of(x) & 0x7fffffff;
  if (ix < 0x3e400000) {
    if ((int)x == 0) return x;
  }
  double z = x * x;
  double v = z * x;
  double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  if (iy == 0) return x + v * (S1 + z * r);
  return x - ((z * (half * y - v * r) - y) - v * S1);
}

/* k_cos.c: cos on [-pi/4, pi/4] */
static double k_cos(double x, double y) {
  const double one = 1.0;
  const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03, C3 = 2.48015872894767294178e-05;
  const double C4 = -2.75573143513906633035e-07, C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
  int32_t ix = hi_of(x) & 0x7fffffff;
  if (ix < 0x3e400000) {
    if ((int)x == 0) return one;
  }
  double z = x * x;
  double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3FD33333) return one - (0.5 * z - (z * r - x * y));
  double qx;
  if (ix > 0x3fe90000) qx = 0.28125;
  else qx = words(ix - 0x00200000, 0);
  double hz = 0.5 * z - qx;
  double a = one - qx;
  return a - (hz - (z * r - x * y));
}

/* e_rem_pio2.c for |x| <= 2^19*pi/2; returns n with x - n*pi/2 = y0 + y1. The caller handles larger |x|. */
static const int32_t npio2_hw[32] = {
    0x3FF921FB, 0x400921FB, 0x4012D97C, 0x401921FB, 0x401F6A7A, 0x4022D97C, 0x4025FDBB, 0x402921FB,
    0x402C463A, 0x402F6A7A, 0x4031475C, 0x4032D97C, 0x40346B9C, 0x4035FDBB, 0x40378FDB, 0x403921FB,
    0x403AB41B, 0x403C463A, 0x403DD85A, 0x403F6A7A, 0x40407E4C, 0x4041475C, 0x4042106C, 0x4042D97C,
    0x4043A28C, 0x40446B9C, 0x404534AC, 0x4045FDBB, 0x4046C6CB, 0x40478FDB, 0x404858EB, 0x404921FB,
};

static int rem_pio2(double x, double* y0, double* y1) {
  const double half = 5.00000000000000000000e-01, invpio2 = 6.36619772367581382433e-01;
  const double pio2_1 = 1.57079632673412561417e+00, pio2_1t = 6.07710050650619224932e-11;
  const double pio2_2 = 6.07710050630396597660e-11, pio2_2t = 2.02226624879595063154e-21;
  const double pio2_3 = 2.02226624871116645580e-21, pio2_3t = 8.47842766036889956997e-32;
  int32_t hx = hi_of(x);
  int32_t ix = hx & 0x7fffffff;
  if (ix <= 0x3fe921fb) {
    *y0 = x;
    *y1 = 0;
    return 0;
  }
  if (ix < 0x4002d97c) {
    double z;
    if (hx > 0) {
      z = x - pio2_1;
      if (ix != 0x3ff921fb) {
        *y0 = z - pio2_1t;
        *y1 = (z - *y0) - pio2_1t;
      } else {
        z -= pio2_2;
        *y0 = z - pio2_2t;
        *y1 = (z - *y0) - pio2_2t;
      }
      return 1;
    }
    z = x + pio2_1;
    if (ix != 0x3ff921fb) {
      *y0 = z + pio2_1t;
      *y1 = (z - *y0) + pio2_1t;
    } else {
      z += pio2_2;
      *y0 = z + pio2_2t;
      *y1 = (z - *y0) + pio2_2t;
    }
    return -1;
  }
  /* medium size: ix <= 0x413921fb (checked by the callers) */
  double t = fabs(x);
  int n = (int)(t * invpio2 + half);
  double fn = (double)n;
  double r = t - fn * pio2_1;
  double w = fn * pio2_1t;
  if (n < 32 && ix != npio2_hw[n - 1]) {
    *y0 = r - w;
  } else {
    int32_t j = ix >> 20;
    *y0 = r - w;
    int32_t i = j - ((hi_of(*y0) >> 20) & 0x7ff);
    if (i > 16) {
      t = r;
      w = fn * pio2_2;
      r = t - w;
      w = fn * pio2_2t - ((t - r) - w);
      *y0 = r - w;
      i = j - ((hi_of(*y0) >> 20) & 0x7ff);
      if (i > 49) {
        t = r;
        w = fn * pio2_3;
        r = t - w;
        w = fn * pio2_3t - ((t - r) - w);
        *y0 = r - w;
      }
    }
  }
  *y1 = (r - *y0) - w;
  if (hx < 0) {
    *y0 = -*y0;
    *y1 = -*y1;
    return -n;
  }
  return n;
}

/* s_sin.c */
double orc_fdl_sin(double x) {
  int32_t ix = hi_of(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) return k_sin(x, 0.0, 0);
  if (ix >= 0x7ff00000) return x - x;
  if (ix > 0x413921fb) return sin(x);
  double y0, y1;
  int n = rem_pio2(x, &y0, &y1);
  switch (n & 3) {
    case 0: return k_sin(y0, y1, 1);
    case 1: return k_cos(y0, y1);
    case 2: return -k_sin(y0, y1, 1);
    default: return -k_cos(y0, y1);
  }
}

/* s_cos.c */
double orc_fdl_cos(double x) {
  int32_t ix = hi_of(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) return k_cos(x, 0.0);
  if (ix >= 0x7ff00000) return x - x;
  if (ix > 0x413921fb) return cos(x);
  double y0, y1;
  int n = rem_pio2(x, &y0, &y1);
  switch (n & 3) {
    case 0: return k_cos(y0, y1);
    case 1: return -k_sin(y0, y1, 1);
    case 2: return -k_cos(y0, y1);
    default: return k_sin(y0, y1, 1);
  }
}

/* e_pow.c (with w_pow.c under _IEEE_LIBM: no wrapper error handling) */
double orc_fdl_pow(double x, double y) {
  static const double bp[2] = {1.0, 1.5};
  static const double dp_h[2] = {0.0, 5.84962487220764160156e-01};
  static const double dp_l[2] = {0.0, 1.35003920212974897128e-08};
  const double zero = 0.0, one = 1.0, two = 2.0, two53 = 9007199254740992.0, huge = 1.0e300, tiny = 1.0e-300;
  const double L1 = 5.99999999999994648725e-01, L2 = 4.28571428578550184252e-01, L3 = 3.33333329818377432918e-01;
  const double L4 = 2.72728123808534006489e-01, L5 = 2.30660745775561754067e-01, L6 = 2.06975017800338417784e-01;
  const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03, P3 = 6.61375632143793436117e-05;
  const double P4 = -1.65339022054652515390e-06, P5 = 4.13813679705723846039e-08;
  const double lg2 = 6.93147180559945286227e-01, lg2_h = 6.93147182464599609375e-01, lg2_l = -1.90465429995776804525e-09;
  const double ovt = 8.0085662595372944372e-0017;
  const double cp = 9.61796693925975554329e-01, cp_h = 9.61796700954437255859e-01, cp_l = -7.02846165095275826516e-09;
  const double ivln2 = 1.44269504088896338700e+00, ivln2_h = 1.44269502162933349609e+00, ivln2_l = 1.92596299112661746887e-08;

  double z, ax, z_h, z_l, p_h, p_l, y1, t1, t2, r, s, t, u, v, w;
  int32_t i, j, k, yisint, n, hx, hy, ix, iy;
  uint32_t lx, ly;

  hx = hi_of(x);
  lx = lo_of(x);
  hy = hi_of(y);
  ly = lo_of(y);
  ix = hx & 0x7fffffff;
  iy = hy & 0x7fffffff;

  if (((uint32_t)iy | ly) == 0) return one;
  if (ix > 0x7ff00000 || ((ix == 0x7ff00000) && (lx != 0)) || iy > 0x7ff00000 || ((iy == 0x7ff00000) && (ly != 0)))
    return x + y;

  /* yisint: 0 not an integer, 1 odd, 2 even (only needed when x < 0) */
  yisint = 0;
  if (hx < 0) {
    if (iy >= 0x43400000) yisint = 2;
    else if (iy >= 0x3ff00000) {
      k = (iy >> 20) - 0x3ff;
      if (k > 20) {
        uint32_t jj = ly >> (52 - k);
        if ((jj << (52 - k)) == ly) yisint = 2 - (int32_t)(jj & 1);
      } else if (ly == 0) {
        uint32_t jj = (uint32_t)iy >> (20 - k);
        if ((jj << (20 - k)) == (uint32_t)iy) yisint = 2 - (int32_t)(jj & 1);
      }
    }
  }

  if (ly == 0) {
    if (iy == 0x7ff00000) {
      if ((((uint32_t)ix - 0x3ff00000u) | lx) == 0) return y - y;
      else if (ix >= 0x3ff00000) return (hy >= 0) ? y : zero;
      else return (hy < 0) ? -y : zero;
    }
    if (iy == 0x3ff00000) {
      if (hy < 0) return one / x;
      return x;
    }
    if (hy == 0x40000000) return x * x;
    if (hy == 0x3fe00000) {
      if (hx >= 0) return sqrt(x);
    }
  }

  ax = fabs(x);
  if (lx == 0) {
    if (ix == 0x7ff00000 || ix == 0 || ix == 0x3ff00000) {
      z = ax;
      if (hy < 0) z = one / z;
      if (hx < 0) {
        if (((ix - 0x3ff00000) | yisint) == 0) {
          z = (z - z) / (z - z);
        } else if (yisint == 1) {
          z = -z;
        }
      }
      return z;
    }
  }

  n = (hx < 0) ? 0 : 1; /* fdlibm's (hx>>31)+1 with an arithmetic shift */

  if ((n | yisint) == 0) return (x - x) / (x - x);

  s = one;
  if ((n | (yisint - 1)) == 0) s = -one;

  if (iy > 0x41e00000) {
    if (iy > 0x43f00000) {
      if (ix <= 0x3fefffff) return (hy < 0) ? huge * huge : tiny * tiny;
      if (ix >= 0x3ff00000) return (hy > 0) ? huge * huge : tiny * tiny;
    }
    if (ix < 0x3fefffff) return (hy < 0) ? s * huge * huge : s * tiny * tiny;
    if (ix > 0x3ff00000) return (hy > 0) ? s * huge * huge : s * tiny * tiny;
    t = ax - one;
    w = (t * t) * (0.5 - t * (0.3333333333333333333333 - t * 0.25));
    u = ivln2_h * t;
    v = t * ivln2_l - w * ivln2;
    t1 = u + v;
    t1 = set_lo(t1, 0);
    t2 = v - (t1 - u);
  } else {
    double ss, s2, s_h, s_l, t_h, t_l;
    n = 0;
    if (ix < 0x00100000) {
      ax *= two53;
      n -= 53;
      ix = hi_of(ax);
    }
    n += ((ix) >> 20) - 0x3ff;
    j = ix & 0x000fffff;
    ix = j | 0x3ff00000;
    if (j <= 0x3988E) k = 0;
    else if (j < 0xBB67A) k = 1;
    else {
      k = 0;
      n += 1;
      ix -= 0x00100000;
    }
    ax = set_hi(ax, ix);

    u = ax - bp[k];
    v = one / (ax + bp[k]);
    ss = u * v;
    s_h = ss;
    s_h = set_lo(s_h, 0);
    t_h = words(((ix >> 1) | 0x20000000) + 0x00080000 + (k << 18), 0);
    t_l = ax - (t_h - bp[k]);
    s_l = v * ((u - s_h * t_h) - s_h * t_l);
    s2 = ss * ss;
    r = s2 * s2 * (L1 + s2 * (L2 + s2 * (L3 + s2 * (L4 + s2 * (L5 + s2 * L6)))));
    r += s_l * (s_h + ss);
    s2 = s_h * s_h;
    t_h = 3.0 + s2 + r;
    t_h = set_lo(t_h, 0);
    t_l = r - ((t_h - 3.0) - s2);
    u = s_h * t_h;
    v = s_l * t_h + t_l * ss;
    p_h = u + v;
    p_h = set_lo(p_h, 0);
    p_l = v - (p_h - u);
    z_h = cp_h * p_h;
    z_l = cp_l * p_h + p_l * cp + dp_l[k];
    t = (double)n;
    t1 = (((z_h + z_l) + dp_h[k]) + t);
    t1 = set_lo(t1, 0);
    t2 = z_l - (((t1 - t) - dp_h[k]) - z_h);
  }

  y1 = y;
  y1 = set_lo(y1, 0);
  p_l = (y - y1) * t1 + y * t2;
  p_h = y1 * t1;
  z = p_l + p_h;
  j = hi_of(z);
  i = (int32_t)lo_of(z);
  if (j >= 0x40900000) {
    if (((j - 0x40900000) | i) != 0) return s * huge * huge;
    if (p_l + ovt > z - p_h) return s * huge * huge;
  } else if ((j & 0x7fffffff) >= 0x4090cc00) {
    if (((uint32_t)j - 0xc090cc00u | (uint32_t)i) != 0) return s * tiny * tiny;
    if (p_l <= z - p_h) return s * tiny * tiny;
  }
  i = j & 0x7fffffff;
  k = (i >> 20) - 0x3ff;
  n = 0;
  if (i > 0x3fe00000) {
    n = j + (0x00100000 >> (k + 1));
    k = ((n & 0x7fffffff) >> 20) - 0x3ff;
    t = words(n & ~(0x000fffff >> k), 0);
    n = ((n & 0x000fffff) | 0x00100000) >> (20 - k);
    if (j < 0) n = -n;
    p_h -= t;
  }
  t = p_l + p_h;
  t = set_lo(t, 0);
  u = t * lg2_h;
  v = (p_l - (t - p_h)) * lg2 + t * lg2_l;
  z = u + v;
  w = v - (z - u);
  t = z * z;
  t1 = z - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
  r = (z * t1) / (t1 - two) - (w + z * w);
  z = one - (r - z);
  j = hi_of(z);
  j = (int32_t)((uint32_t)j + ((uint32_t)n << 20));
  if ((j >> 20) <= 0) z = orc_fdl_scalbn(z, n);
  else z = set_hi(z, (int32_t)((uint32_t)hi_of(z) + ((uint32_t)n << 20)));
  return s * z;
}
