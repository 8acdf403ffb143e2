// pixel.hip -- pointing transform + HEALPix RING ang2pix for every sample.
//
// Compiled with -ffp-contract=off: every expression below keeps the operand
// order of src/sanepic.cc (pointingshift 3412-3476, the pointing loop
// 729-773, the pixelisation loop 848-878 and ang2pix_ring 2764-2843), so the
// only difference from the reference's GNU build is the device libm (cos,
// sin, acos, atan2), which differs from glibc by at most a few ulp.  Each
// floor() argument and branch test whose distance to its decision boundary
// is below that error bound is reported as "ambiguous"; the host resolves
// those samples with the platform libm (pixel_host.cpp).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace {

constexpr double kD2PI = 6.2831853071795864769252867665590057683943387987502;

struct Guard {
  double eps;  // absolute margin on floor() arguments
  bool amb;
  __device__ void arg(double a) {
    if (fabs(a - rint(a)) < eps) amb = true;
  }
};

// Returns the RING pixel; sets g.amb when a device-vs-glibc ulp difference
// could change it.
__device__ int ang2pix_dev(int nside, double theta, double phi, Guard &g) {
  const double piover2 = 0.5 * M_PI;
  const double twopi = 2.0 * M_PI;
  const double z0 = 2.0 / 3.0;
  const int nl2 = 2 * nside, nl4 = 4 * nside;
  const int ncap = nl2 * (nside - 1);
  const int npix = 12 * nside * nside;
  double z = cos(theta);
  double za = fabs(z);
  if (phi >= twopi) phi = phi - twopi;
  if (phi < 0.) phi = phi + twopi;
  double tt = phi / piover2;
  if (tt < g.eps || tt > 4.0 - g.eps) g.amb = true;
  if (fabs(za - z0) < 1e-12) g.amb = true;
  int jp, jm, ir, ip, ipix1;
  if (za <= z0) {
    double a1 = (double)nside * (0.5 + tt - z * 0.75);
    double a2 = (double)nside * (0.5 + tt + z * 0.75);
    g.arg(a1);
    g.arg(a2);
    jp = (int)floor(a1);
    jm = (int)floor(a2);
    ir = nside + 1 + jp - jm;
    int kshift = ((ir & 1) == 0) ? 1 : 0;
    ip = (int)floor((double)((jp + jm - nside + kshift + 1) / 2)) + 1;
    if (ip > nl4) ip = ip - nl4;
    ipix1 = ncap + nl4 * (ir - 1) + ip;
  } else {
    if (1.0 - za < 1e-8) g.amb = true;  // near a pole acos/cos errors grow
    g.arg(tt);
    double tp = tt - floor(tt);
    double tmp = sqrt(3. * (1. - za));
    double a1 = (double)nside * tp * tmp;
    double a2 = (double)nside * (1. - tp) * tmp;
    g.arg(a1);
    g.arg(a2);
    jp = (int)floor(a1);
    jm = (int)floor(a2);
    ir = jp + jm + 1;
    double a3 = tt * ir;
    g.arg(a3);
    ip = (int)floor(a3) + 1;
    if (ip > 4 * ir) ip = ip - 4 * ir;
    ipix1 = 2 * ir * (ir - 1) + ip;
    if (z <= 0.) ipix1 = npix - 2 * ir * (ir + 1) + ip;
  }
  return ipix1 - 1;
}

__device__ __forceinline__ bool isnan_(double x) { return !(x > 0) && !(x <= 0.0); }

template <bool HWP>
__global__ __launch_bounds__(256) void k_pixelize(PixelArgs a) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.ntot) return;
  const int64_t ns = a.ns;
  const int64_t t = g % ns;
  const int64_t seg = (g / ns) % a.nseg;
  const int det = (int)(g / (ns * a.nseg));
  const int64_t bs = seg * ns + t;
  const DetConst dc = a.det[det];
  // psi NaN fix of 722-727: the last non-NaN value at or before t
  double psv = a.psi[bs];
  if (isnan_(psv)) {
    int64_t k = t - 1;
    while (k >= 0 && isnan_(a.psi[seg * ns + k])) k--;
    if (k >= 0) psv = a.psi[seg * ns + k];
  }
  const double dec = a.dec[bs];
  const double ra = a.ra[bs];
  // pointingshift (3419-3428) with nn = 1; sin/cos of theta_ from the host
  const double thetap = kD2PI / 4.0 - dec;
  const double phip = ra;
  const double cps = cos(dc.ph_ + psv);
  const double theta_o = acos(sin(thetap) * dc.sin_th * cps + cos(thetap) * dc.cos_th);
  const double phi_o =
      atan2(-dc.sin_th * sin(dc.ph_ + psv), (-cos(thetap) * dc.sin_th * cps + sin(thetap) * dc.cos_th)) +
      phip;
  // psi (3435-3472)
  double drp0 = cos(psv) * dc.cos_th;
  double drp1 = sin(psv) * dc.cos_th;
  double drp2 = dc.drp2;
  double R00 = -cos(thetap), R02 = sin(thetap), R20 = sin(thetap), R22 = cos(thetap);
  double dr0 = 0.0, dr1 = 0.0, dr2 = 0.0;
  dr0 += R00 * drp0;
  dr0 += 0.0 * drp1;
  dr0 += R02 * drp2;
  dr1 += 0.0 * drp0;
  dr1 += -1.0 * drp1;
  dr1 += 0.0 * drp2;
  dr2 += R20 * drp0;
  dr2 += 0.0 * drp1;
  dr2 += R22 * drp2;
  double eph0 = -sin(phi_o - phip);
  double eph1 = cos(phi_o - phip);
  double rrr = 0.0, aaa = 0.0;
  rrr += dr0 * dr0;
  aaa += dr0 * eph0;
  rrr += dr1 * dr1;
  aaa += dr1 * eph1;
  rrr += dr2 * dr2;
  aaa += dr2 * 0.0;
  aaa = aaa / sqrt(rrr);
  if (aaa > 1.0) aaa = 1.0;
  if (aaa < -1.0) aaa = -1.0;
  double pt = acos(aaa);
  if (dr2 < 0) pt = -pt;
  pt = (fmod(pt + kD2PI / 4.0 + 2.0 * kD2PI, kD2PI)) - kD2PI / 2.0;
  double psip = pt + dc.psidet;
  const double psiFP = psip;
  if (HWP) psip = -(pt + dc.psidet) + 2.0 * a.omega * (double)(t + ns * seg);
  // pixelisation (855-860)
  const double ra_d = phi_o;
  const double dec_d = kD2PI / 4.0 - theta_o;
  const double theta = kD2PI / 4.0 - dec_d;
  const double phi = fmod(ra_d + kD2PI, kD2PI);
  Guard gd{a.eps, false};
  int pix = -1;
  bool bad = isnan_(theta) || isnan_(phi) || theta < 0. || theta > M_PI;
  if (!bad) {
    pix = ang2pix_dev(a.nside, theta, phi, gd);
    if (pix < 0 || pix > 12 * a.nside * a.nside - 1) bad = true;
  }
  const double c2 = cos(2.0 * (psip + 0.0));
  const double s2 = sin(2.0 * (psip + 0.0));
  a.cs[g] = make_double2(c2, s2);
  if (HWP) a.csfp[g] = make_double2(cos(2.0 * (psiFP + 0.0)), sin(2.0 * (psiFP + 0.0)));
  if (bad || gd.amb) {
    unsigned long long slot = atomicAdd(a.nflag, 1ULL);
    if (slot < (unsigned long long)a.flag_cap) {
      FlagRec r;
      r.g = g;
      r.ra = ra;
      r.dec = dec;
      r.psi = psv;
      r.bad = bad ? 1 : 0;
      a.flags[slot] = r;
    }
    pix = bad ? -1 : pix;
  }
  a.pix[g] = pix;
}

}  // namespace

int launch_pixelize(const PixelArgs &a, hipStream_t s) {
  const int64_t nb = (a.ntot + 255) / 256;
  if (a.hwp)
    hipLaunchKernelGGL(k_pixelize<true>, dim3((unsigned)nb), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_pixelize<false>, dim3((unsigned)nb), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
