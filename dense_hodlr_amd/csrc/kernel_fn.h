/*
 * Kernel-entry functions K(i,j) for the typed kernel families of the HODLR
 * engine.  One definition serves host code and device code, so an entry
 * evaluated inside a GPU kernel is bit-identical to the same entry evaluated
 * on the host: every operation below is an IEEE-754 correctly rounded
 * operation (add, mul, div, sqrt, floor, ldexp) performed in a fixed order,
 * with floating-point contraction switched off.
 *
 * The reference evaluates entries through a user callback
 * double(*)(int i,int j,void*) (include/kernel.hpp:27, src/kernel.cpp:19-25);
 * a host function pointer cannot run on the GPU, so the engine takes a typed
 * descriptor instead and evaluates the family here.
 */
#ifndef DENSE_HODLR_KERNEL_FN_H
#define DENSE_HODLR_KERNEL_FN_H

#include <math.h>

#if defined(__HIPCC__)
#define HODLR_HD __host__ __device__ inline
#else
#define HODLR_HD inline
#endif

#if defined(__clang__)
#define HODLR_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define HODLR_NO_CONTRACT
#endif

/* Smallest magnitude an ACA pivot candidate may have, 2^-511: its square and
 * the square of its reciprocal are normal FP64 numbers.  The reference
 * (lowRank.cpp:170-227, chooseNextRowCol :291-298) accepts any value above
 * minPivot (an int, normally 0), so on kernels whose far entries underflow
 * (a squared exponential with l = 0.01 on the unit interval) it divides by
 * subnormal pivots and returns non-finite factors.  Candidates below this
 * floor count as zero here, in the oracle and in the GPU kernels alike;
 * inputs whose candidates all exceed it behave exactly as in the reference. */
#define HODLR_PIVOT_FLOOR 0x1p-511

/* Kernel families.  r is the Euclidean distance, l the length scale. */
enum {
  HODLR_KF_GAUSSIAN = 0,    /* exp(-r^2 / l^2)                      */
  HODLR_KF_MATERN32 = 1,    /* (1 + sqrt(3) r / l) exp(-sqrt(3) r/l) */
  HODLR_KF_SE = 2,          /* exp(-r^2 / (2 l^2))                  */
  HODLR_KF_IMQ = 3,         /* 1 / sqrt(1 + r^2 / l^2)              */
  HODLR_KF_QUADRATIC = 4,   /* 1 + r^2 / l^2                        */
  HODLR_KF_EXPONENTIAL = 5  /* exp(-r / l)                          */
};

/* Diagonal handling: 0 adds diag to f(0); 1 replaces K(i,i) by diag (the
 * convention of the reference's 1D test matrices,
 * src/helperFunctions.cpp:81-82). */
enum { HODLR_DIAG_ADD = 0, HODLR_DIAG_REPLACE = 1 };

/* Parameters in the form the evaluator uses.  Coordinates are stored
 * structure-of-arrays (x[i], y[i], z[i]). */
struct HodlrKernelParams {
  int family;
  int dim;
  int diag_mode;
  int pad_;
  double inv_l;    /* 1 / l                 */
  double inv_l2;   /* 1 / (l * l)           */
  double s3_inv_l; /* sqrt(3) / l           */
  double diag;     /* diagonal shift/value  */
  const double* x;
  const double* y;
  const double* z;
};

/*
 * exp(x) by Cody-Waite reduction x = n ln2 + r, |r| <= ln2/2, and the Taylor
 * series of e^r to degree 13 in Horner form (truncation < 5e-18 relative),
 * then scaling by 2^n with ldexp.  ln2 is split into a 32-bit head, so n*LN2_HI
 * is exact for |n| < 2^11, and a tail.  Accuracy is within a few ulp of the
 * correctly rounded result over the whole double range; being built only from
 * correctly rounded operations it returns the same bits on host and device.
 */
HODLR_HD double hodlr_exp(double x) {
  HODLR_NO_CONTRACT
  if (x != x) return x;
  if (x < -746.0) return 0.0;
  if (x > 709.78) return HUGE_VAL;
  const double INV_LN2 = 1.4426950408889634074;
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  double n = floor(x * INV_LN2 + 0.5);
  double r = (x - n * LN2_HI) - n * LN2_LO;
  double p = 1.0 / 6227020800.0;          /* 1/13! */
  p = p * r + 1.0 / 479001600.0;          /* 1/12! */
  p = p * r + 1.0 / 39916800.0;           /* 1/11! */
  p = p * r + 1.0 / 3628800.0;            /* 1/10! */
  p = p * r + 1.0 / 362880.0;             /* 1/9!  */
  p = p * r + 1.0 / 40320.0;              /* 1/8!  */
  p = p * r + 1.0 / 5040.0;               /* 1/7!  */
  p = p * r + 1.0 / 720.0;                /* 1/6!  */
  p = p * r + 1.0 / 120.0;                /* 1/5!  */
  p = p * r + 1.0 / 24.0;                 /* 1/4!  */
  p = p * r + 1.0 / 6.0;                  /* 1/3!  */
  p = p * r + 0.5;                        /* 1/2!  */
  p = p * r + 1.0;                        /* 1/1!  */
  p = p * r + 1.0;                        /* 1/0!  */
  return ldexp(p, (int)n);
}

/* Squared distance between points i and j, summed over coordinates in order. */
HODLR_HD double hodlr_dist2(const HodlrKernelParams& p, int i, int j) {
  HODLR_NO_CONTRACT
  double d = p.x[i] - p.x[j];
  double d2 = d * d;
  if (p.dim > 1) {
    d = p.y[i] - p.y[j];
    d2 = d2 + d * d;
  }
  if (p.dim > 2) {
    d = p.z[i] - p.z[j];
    d2 = d2 + d * d;
  }
  return d2;
}

/* f(d2) for the family, without the diagonal term. */
HODLR_HD double hodlr_radial(const HodlrKernelParams& p, double d2) {
  HODLR_NO_CONTRACT
  switch (p.family) {
    case HODLR_KF_GAUSSIAN:
      return hodlr_exp(-(d2 * p.inv_l2));
    case HODLR_KF_SE:
      return hodlr_exp(-(0.5 * (d2 * p.inv_l2)));
    case HODLR_KF_MATERN32: {
      double t = p.s3_inv_l * sqrt(d2);
      return (1.0 + t) * hodlr_exp(-t);
    }
    case HODLR_KF_IMQ:
      return 1.0 / sqrt(1.0 + d2 * p.inv_l2);
    case HODLR_KF_QUADRATIC:
      return 1.0 + d2 * p.inv_l2;
    case HODLR_KF_EXPONENTIAL:
      return hodlr_exp(-(sqrt(d2) * p.inv_l));
    default:
      return 0.0;
  }
}

/* Entry K(i,j) with global indices i, j. */
HODLR_HD double hodlr_kernel_entry(const HodlrKernelParams& p, int i, int j) {
  HODLR_NO_CONTRACT
  if (i == j && p.diag_mode == HODLR_DIAG_REPLACE) return p.diag;
  double v = hodlr_radial(p, hodlr_dist2(p, i, j));
  if (i == j) v = v + p.diag;
  return v;
}

/* Fill the derived constants from a family, dimension and length scale. */
HODLR_HD void hodlr_kernel_params_init(HodlrKernelParams* p, int family, int dim,
                                       double length_scale, double diag,
                                       int diag_mode, const double* x,
                                       const double* y, const double* z) {
  HODLR_NO_CONTRACT
  p->family = family;
  p->dim = dim;
  p->diag_mode = diag_mode;
  p->pad_ = 0;
  p->inv_l = 1.0 / length_scale;
  p->inv_l2 = 1.0 / (length_scale * length_scale);
  p->s3_inv_l = 1.7320508075688772935 / length_scale;
  p->diag = diag;
  p->x = x;
  p->y = y;
  p->z = z;
}

#endif /* DENSE_HODLR_KERNEL_FN_H */
