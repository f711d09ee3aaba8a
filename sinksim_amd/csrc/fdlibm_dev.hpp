// fdlibm 5.3 sin, cos, pow and scalbn as host+device functions for gfx950, so that the step's transcendental
// calls give the same bits on the GPU as the reference's portable numerics on any CPU
// (reference: engine/include/sinksim/kernel/fdlibm_portable.hpp:20-435, itself netlib fdlibm 5.3 from
// third_party/fdlibm/src/{s_sin,s_cos,k_sin,k_cos,e_rem_pio2,e_pow,s_scalbn,s_copysign}.c).
//
// The algorithm and its constants are fdlibm's (Sun Microsystems; "Permission to use, copy, modify, and
// distribute this software is freely granted, provided that this notice is preserved."). The code is a
// restatement for the GPU: bit access through __builtin_bit_cast, and sin and cos folded into one routine that
// differs only in the quadrant, so that the lanes computing the pose's four trigonometric values follow one
// instruction stream. Arguments of sin/cos above 2^19*pi/2 would need fdlibm's large reduction and use the
// platform function, as the reference does (fdlibm_portable.hpp:224,239); the step clamps its angles to 1.4 rad.
//
// Compile with -ffp-contract=off: every product and sum here must round on its own.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ssa::fdl {

__host__ __device__ inline int32_t hiw(double x) { return static_cast<int32_t>(__builtin_bit_cast(uint64_t, x) >> 32); }
__host__ __device__ inline uint32_t low(double x) { return static_cast<uint32_t>(__builtin_bit_cast(uint64_t, x)); }
__host__ __device__ inline double mk(int32_t h, uint32_t l) {
  return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(h)) << 32) | static_cast<uint64_t>(l));
}
__host__ __device__ inline double with_low(double x, uint32_t l) { return mk(hiw(x), l); }
__host__ __device__ inline double with_hi(double x, int32_t h) { return mk(h, low(x)); }

__host__ __device__ inline double fabs_(double x) { return mk(hiw(x) & 0x7fffffff, low(x)); }

__host__ __device__ inline double copysign_(double x, double y) {
  return with_hi(x, static_cast<int32_t>((static_cast<uint32_t>(hiw(x)) & 0x7fffffffu) | (static_cast<uint32_t>(hiw(y)) & 0x80000000u)));
}

// s_scalbn.c
__host__ __device__ inline double scalbn(double x, int n) {
  constexpr double two54 = 1.80143985094819840000e+16, twom54 = 5.55111512312578270212e-17;
  constexpr double huge = 1.0e+300, tiny = 1.0e-300;
  int32_t hx = hiw(x);
  const uint32_t lx = low(x);
  int32_t k = (hx & 0x7ff00000) >> 20;
  if (k == 0) {
    if ((lx | static_cast<uint32_t>(hx & 0x7fffffff)) == 0) return x;
    x *= two54;
    hx = hiw(x);
    k = ((hx & 0x7ff00000) >> 20) - 54;
    if (n < -50000) return tiny * x;
  }
  if (k == 0x7ff) return x + x;
  k = k + n;
  if (k > 0x7fe) return huge * copysign_(huge, x);
  if (k > 0) return with_hi(x, static_cast<int32_t>((static_cast<uint32_t>(hx) & 0x800fffffu) | (static_cast<uint32_t>(k) << 20)));
  if (k <= -54) {
    if (n > 50000) return huge * copysign_(huge, x);
    return tiny * copysign_(tiny, x);
  }
  k += 54;
  x = with_hi(x, static_cast<int32_t>((static_cast<uint32_t>(hx) & 0x800fffffu) | (static_cast<uint32_t>(k) << 20)));
  return x * twom54;
}

// k_sin.c
__host__ __device__ inline double ksin(double x, double y, int iy) {
  constexpr double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03, S3 = -1.98412698298579493134e-04,
                   S4 = 2.75573137070700676789e-06, S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
  const int32_t ix = hiw(x) & 0x7fffffff;
  if (ix < 0x3e400000 && static_cast<int>(x) == 0) return x;
  const double z = x * x;
  const double v = z * x;
  const double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  if (iy == 0) return x + v * (S1 + z * r);
  return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

// k_cos.c
__host__ __device__ inline double kcos(double x, double y) {
  constexpr double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03, C3 = 2.48015872894767294178e-05,
                   C4 = -2.75573143513906633035e-07, C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
  const int32_t ix = hiw(x) & 0x7fffffff;
  if (ix < 0x3e400000 && static_cast<int>(x) == 0) return 1.0;
  const double z = x * x;
  const double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3FD33333) return 1.0 - (0.5 * z - (z * r - x * y));
  const double qx = ix > 0x3fe90000 ? 0.28125 : mk(ix - 0x00200000, 0u);
  const double hz = 0.5 * z - qx;
  const double a = 1.0 - qx;
  return a - (hz - (z * r - x * y));
}

// e_rem_pio2.c for |x| <= 2^19*pi/2 (ix <= 0x413921fb, checked by the caller): x - n*pi/2 = y0 + y1
__host__ __device__ inline int rem_pio2(double x, double& y0, double& y1) {
  constexpr double invpio2 = 6.36619772367581382433e-01;
  constexpr double pio2_1 = 1.57079632673412561417e+00, pio2_1t = 6.07710050650619224932e-11;
  constexpr double pio2_2 = 6.07710050630396597660e-11, pio2_2t = 2.02226624879595063154e-21;
  constexpr double pio2_3 = 2.02226624871116645580e-21, pio2_3t = 8.47842766036889956997e-32;
  const int32_t hx = hiw(x);
  const int32_t ix = hx & 0x7fffffff;
  if (ix < 0x4002d97c) {  // |x| < 3pi/4: n = +-1
    const bool nearHalfPi = ix == 0x3ff921fb;
    if (hx > 0) {
      double z = x - pio2_1;
      if (!nearHalfPi) {
        y0 = z - pio2_1t;
        y1 = (z - y0) - pio2_1t;
      } else {
        z -= pio2_2;
        y0 = z - pio2_2t;
        y1 = (z - y0) - pio2_2t;
      }
      return 1;
    }
    double z = x + pio2_1;
    if (!nearHalfPi) {
      y0 = z + pio2_1t;
      y1 = (z - y0) + pio2_1t;
    } else {
      z += pio2_2;
      y0 = z + pio2_2t;
      y1 = (z - y0) + pio2_2t;
    }
    return -1;
  }
  // high words of n*pi/2 for n = 1..32: below 32 a differing high word proves there is no cancellation
  constexpr int32_t npio2_hw[32] = {
      0x3FF921FB, 0x400921FB, 0x4012D97C, 0x401921FB, 0x401F6A7A, 0x4022D97C, 0x4025FDBB, 0x402921FB,
      0x402C463A, 0x402F6A7A, 0x4031475C, 0x4032D97C, 0x40346B9C, 0x4035FDBB, 0x40378FDB, 0x403921FB,
      0x403AB41B, 0x403C463A, 0x403DD85A, 0x403F6A7A, 0x40407E4C, 0x4041475C, 0x4042106C, 0x4042D97C,
      0x4043A28C, 0x40446B9C, 0x404534AC, 0x4045FDBB, 0x4046C6CB, 0x40478FDB, 0x404858EB, 0x404921FB};
  double t = fabs_(x);
  const int n = static_cast<int>(t * invpio2 + 0.5);
  const double fn = static_cast<double>(n);
  double r = t - fn * pio2_1;
  double w = fn * pio2_1t;
  y0 = r - w;
  if (!(n < 32 && ix != npio2_hw[n - 1])) {
    const int32_t j = ix >> 20;
    int32_t i = j - ((hiw(y0) >> 20) & 0x7ff);
    if (i > 16) {
      t = r;
      w = fn * pio2_2;
      r = t - w;
      w = fn * pio2_2t - ((t - r) - w);
      y0 = r - w;
      i = j - ((hiw(y0) >> 20) & 0x7ff);
      if (i > 49) {
        t = r;
        w = fn * pio2_3;
        r = t - w;
        w = fn * pio2_3t - ((t - r) - w);
        y0 = r - w;
      }
    }
  }
  y1 = (r - y0) - w;
  if (hx < 0) {
    y0 = -y0;
    y1 = -y1;
    return -n;
  }
  return n;
}

// s_sin.c / s_cos.c in one routine: cos(x) is sin's quadrant table shifted by one.
__host__ __device__ inline double sin_or_cos(double x, int wantCos) {
  const int32_t ix = hiw(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) return wantCos ? kcos(x, 0.0) : ksin(x, 0.0, 0);
  if (ix >= 0x7ff00000) return x - x;
  if (ix > 0x413921fb) {
#if defined(__HIP_DEVICE_COMPILE__)
    return wantCos ? ::cos(x) : ::sin(x);
#else
    return wantCos ? __builtin_cos(x) : __builtin_sin(x);
#endif
  }
  double y0, y1;
  const int n = rem_pio2(x, y0, y1);
  switch ((n + wantCos) & 3) {
    case 0: return ksin(y0, y1, 1);
    case 1: return kcos(y0, y1);
    case 2: return -ksin(y0, y1, 1);
    default: return -kcos(y0, y1);
  }
}

__host__ __device__ inline double sin(double x) { return sin_or_cos(x, 0); }
__host__ __device__ inline double cos(double x) { return sin_or_cos(x, 1); }

// e_pow.c (the _IEEE_LIBM build: no wrapper error handling)
__host__ __device__ inline double pow(double x, double y) {
  constexpr double dp_h1 = 5.84962487220764160156e-01, dp_l1 = 1.35003920212974897128e-08;
  constexpr double two53 = 9007199254740992.0, huge = 1.0e300, tiny = 1.0e-300;
  constexpr double L1 = 5.99999999999994648725e-01, L2 = 4.28571428578550184252e-01, L3 = 3.33333329818377432918e-01,
                   L4 = 2.72728123808534006489e-01, L5 = 2.30660745775561754067e-01, L6 = 2.06975017800338417784e-01;
  constexpr double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03, P3 = 6.61375632143793436117e-05,
                   P4 = -1.65339022054652515390e-06, P5 = 4.13813679705723846039e-08;
  constexpr double lg2 = 6.93147180559945286227e-01, lg2_h = 6.93147182464599609375e-01, lg2_l = -1.90465429995776804525e-09;
  constexpr double ovt = 8.0085662595372944372e-0017;
  constexpr double cp = 9.61796693925975554329e-01, cp_h = 9.61796700954437255859e-01, cp_l = -7.02846165095275826516e-09;
  constexpr double ivln2 = 1.44269504088896338700e+00, ivln2_h = 1.44269502162933349609e+00, ivln2_l = 1.92596299112661746887e-08;

  const int32_t hx = hiw(x), hy = hiw(y);
  const uint32_t lx = low(x), ly = low(y);
  int32_t ix = hx & 0x7fffffff;
  const int32_t iy = hy & 0x7fffffff;

  if ((static_cast<uint32_t>(iy) | ly) == 0) return 1.0;  // x**0 = 1
  if (ix > 0x7ff00000 || (ix == 0x7ff00000 && lx != 0) || iy > 0x7ff00000 || (iy == 0x7ff00000 && ly != 0)) return x + y;

  // y's integer class when x < 0: 0 not an integer, 1 odd, 2 even
  int32_t yisint = 0;
  if (hx < 0) {
    if (iy >= 0x43400000) {
      yisint = 2;
    } else if (iy >= 0x3ff00000) {
      const int32_t e = (iy >> 20) - 0x3ff;
      if (e > 20) {
        const uint32_t jj = ly >> (52 - e);
        if ((jj << (52 - e)) == ly) yisint = 2 - static_cast<int32_t>(jj & 1u);
      } else if (ly == 0) {
        const uint32_t jj = static_cast<uint32_t>(iy) >> (20 - e);
        if ((jj << (20 - e)) == static_cast<uint32_t>(iy)) yisint = 2 - static_cast<int32_t>(jj & 1u);
      }
    }
  }

  if (ly == 0) {
    if (iy == 0x7ff00000) {  // y = +-inf
      if (((static_cast<uint32_t>(ix) - 0x3ff00000u) | lx) == 0) return y - y;
      if (ix >= 0x3ff00000) return hy >= 0 ? y : 0.0;
      return hy < 0 ? -y : 0.0;
    }
    if (iy == 0x3ff00000) return hy < 0 ? 1.0 / x : x;
    if (hy == 0x40000000) return x * x;
    if (hy == 0x3fe00000 && hx >= 0) return __builtin_sqrt(x);
  }

  double ax = fabs_(x);
  if (lx == 0 && (ix == 0x7ff00000 || ix == 0 || ix == 0x3ff00000)) {  // x = +-0, +-inf, +-1
    double z = ax;
    if (hy < 0) z = 1.0 / z;
    if (hx < 0) {
      if (((ix - 0x3ff00000) | yisint) == 0) z = (z - z) / (z - z);
      else if (yisint == 1) z = -z;
    }
    return z;
  }

  const int32_t xpos = hx < 0 ? 0 : 1;
  if ((xpos | yisint) == 0) return (x - x) / (x - x);  // (x<0)**(non-int)
  double s = 1.0;
  if ((xpos | (yisint - 1)) == 0) s = -1.0;  // (-ve)**(odd int)

  double t1, t2;
  if (iy > 0x41e00000) {  // |y| > 2**31
    if (iy > 0x43f00000) {
      if (ix <= 0x3fefffff) return hy < 0 ? huge * huge : tiny * tiny;
      if (ix >= 0x3ff00000) return hy > 0 ? huge * huge : tiny * tiny;
    }
    if (ix < 0x3fefffff) return hy < 0 ? s * huge * huge : s * tiny * tiny;
    if (ix > 0x3ff00000) return hy > 0 ? s * huge * huge : s * tiny * tiny;
    const double t = ax - 1.0;
    const double w = (t * t) * (0.5 - t * (0.3333333333333333333333 - t * 0.25));
    const double u = ivln2_h * t;
    const double v = t * ivln2_l - w * ivln2;
    t1 = with_low(u + v, 0u);
    t2 = v - (t1 - u);
  } else {
    int32_t n = 0;
    if (ix < 0x00100000) {  // subnormal x
      ax *= two53;
      n -= 53;
      ix = hiw(ax);
    }
    n += (ix >> 20) - 0x3ff;
    const int32_t j = ix & 0x000fffff;
    ix = j | 0x3ff00000;
    int32_t k;
    if (j <= 0x3988E) {
      k = 0;
    } else if (j < 0xBB67A) {
      k = 1;
    } else {
      k = 0;
      n += 1;
      ix -= 0x00100000;
    }
    ax = with_hi(ax, ix);
    const double bp = k ? 1.5 : 1.0, dph = k ? dp_h1 : 0.0, dpl = k ? dp_l1 : 0.0;
    double u = ax - bp;
    double v = 1.0 / (ax + bp);
    const double ss = u * v;
    const double s_h = with_low(ss, 0u);
    double t_h = mk(((ix >> 1) | 0x20000000) + 0x00080000 + (k << 18), 0u);
    double t_l = ax - (t_h - bp);
    const double s_l = v * ((u - s_h * t_h) - s_h * t_l);
    double s2 = ss * ss;
    double r = s2 * s2 * (L1 + s2 * (L2 + s2 * (L3 + s2 * (L4 + s2 * (L5 + s2 * L6)))));
    r += s_l * (s_h + ss);
    s2 = s_h * s_h;
    t_h = with_low(3.0 + s2 + r, 0u);
    t_l = r - ((t_h - 3.0) - s2);
    u = s_h * t_h;
    v = s_l * t_h + t_l * ss;
    const double p_h = with_low(u + v, 0u);
    const double p_l = v - (p_h - u);
    const double z_h = cp_h * p_h;
    const double z_l = cp_l * p_h + p_l * cp + dpl;
    const double t = static_cast<double>(n);
    t1 = with_low(((z_h + z_l) + dph) + t, 0u);
    t2 = z_l - (((t1 - t) - dph) - z_h);
  }

  // (y1 + y2) * (t1 + t2) with y split at 32 bits
  const double y1 = with_low(y, 0u);
  double p_l = (y - y1) * t1 + y * t2;
  double p_h = y1 * t1;
  double z = p_l + p_h;
  int32_t j = hiw(z);
  const int32_t i = static_cast<int32_t>(low(z));
  if (j >= 0x40900000) {  // z >= 1024
    if (((j - 0x40900000) | i) != 0) return s * huge * huge;
    if (p_l + ovt > z - p_h) return s * huge * huge;
  } else if ((j & 0x7fffffff) >= 0x4090cc00) {  // z <= -1075
    if (((static_cast<uint32_t>(j) - 0xc090cc00u) | static_cast<uint32_t>(i)) != 0) return s * tiny * tiny;
    if (p_l <= z - p_h) return s * tiny * tiny;
  }
  // 2**(p_h + p_l)
  const int32_t ia = j & 0x7fffffff;
  int32_t e = (ia >> 20) - 0x3ff;
  int32_t n = 0;
  if (ia > 0x3fe00000) {  // |z| > 0.5: n = [z + 0.5]
    n = j + (0x00100000 >> (e + 1));
    e = ((n & 0x7fffffff) >> 20) - 0x3ff;
    const double t = mk(n & ~(0x000fffff >> e), 0u);
    n = ((n & 0x000fffff) | 0x00100000) >> (20 - e);
    if (j < 0) n = -n;
    p_h -= t;
  }
  const double t = with_low(p_l + p_h, 0u);
  const double u = t * lg2_h;
  const double v = (p_l - (t - p_h)) * lg2 + t * lg2_l;
  z = u + v;
  const double w = v - (z - u);
  const double tt = z * z;
  const double tq = z - tt * (P1 + tt * (P2 + tt * (P3 + tt * (P4 + tt * P5))));
  const double r = (z * tq) / (tq - 2.0) - (w + z * w);
  z = 1.0 - (r - z);
  j = static_cast<int32_t>(static_cast<uint32_t>(hiw(z)) + (static_cast<uint32_t>(n) << 20));
  if ((j >> 20) <= 0) z = scalbn(z, n);  // subnormal result
  else z = with_hi(z, j);
  return s * z;
}

}  // namespace ssa::fdl
