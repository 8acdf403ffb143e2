// pixel_host.cpp -- the reference's pixel chain evaluated with the platform
// libm, for the few samples the GPU reports as ambiguous (a floor() argument
// or branch test within the device libm's error bound of its boundary).
//
// Built with g++ -ffp-contract=off and no fast-math: the operation sequence
// is that of src/sanepic.cc (pointingshift 3419-3428, pointing loop
// 737-747, pixelisation 855-860, ang2pix_ring 2794-2842), so the result is
// the reference's pixel whenever libm is glibc.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>
#include <cstdint>

#include "pixel_host.h"

namespace {
constexpr double kD2PI = 6.2831853071795864769252867665590057683943387987502;

long ring_pixel(long nside, double theta, double phi) {
  const double piover2 = 0.5 * M_PI, twopi = 2.0 * M_PI, z0 = 2.0 / 3.0;
  if (nside < 1 || nside > 8192) return -1;
  if (theta < 0. || theta > M_PI) return -1;
  const int nl2 = 2 * (int)nside, nl4 = 4 * (int)nside;
  const int ncap = nl2 * ((int)nside - 1);
  const int npix = 12 * (int)nside * (int)nside;
  const double z = std::cos(theta);
  const double za = std::fabs(z);
  if (phi >= twopi) phi = phi - twopi;
  if (phi < 0.) phi = phi + twopi;
  const double tt = phi / piover2;
  int jp, jm, ir, ip, ipix1;
  if (za <= z0) {
    jp = (int)std::floor(nside * (0.5 + tt - z * 0.75));
    jm = (int)std::floor(nside * (0.5 + tt + z * 0.75));
    ir = (int)nside + 1 + jp - jm;
    const int kshift = ((ir & 1) == 0) ? 1 : 0;
    ip = (int)std::floor((double)((jp + jm - (int)nside + kshift + 1) / 2)) + 1;
    if (ip > nl4) ip = ip - nl4;
    ipix1 = ncap + nl4 * (ir - 1) + ip;
  } else {
    const double tp = tt - std::floor(tt);
    const double tmp = std::sqrt(3. * (1. - za));
    jp = (int)std::floor(nside * tp * tmp);
    jm = (int)std::floor(nside * (1. - tp) * tmp);
    ir = jp + jm + 1;
    ip = (int)std::floor(tt * ir) + 1;
    if (ip > 4 * ir) ip = ip - 4 * ir;
    ipix1 = 2 * ir * (ir - 1) + ip;
    if (z <= 0.) ipix1 = npix - 2 * ir * (ir + 1) + ip;
  }
  return (long)ipix1 - 1;
}
}  // namespace

long host_sample_pixel(long nside, double offx, double offy, double ra, double dec, double psi) {
  const double x = offx, y = offy;
  const double th_ = std::asin(std::sqrt(x * x + y * y));
  const double ph_ = std::atan2(y, x);
  const double thetap = kD2PI / 4.0 - dec;
  const double phip = ra;
  const double theta_o =
      std::acos(std::sin(thetap) * std::sin(th_) * std::cos(ph_ + psi) + std::cos(thetap) * std::cos(th_));
  const double phi_o = std::atan2(-std::sin(th_) * std::sin(ph_ + psi),
                                  (-std::cos(thetap) * std::sin(th_) * std::cos(ph_ + psi) +
                                   std::sin(thetap) * std::cos(th_))) +
                       phip;
  const double ra_d = phi_o;
  const double dec_d = kD2PI / 4.0 - theta_o;
  const double theta = kD2PI / 4.0 - dec_d;
  const double phi = std::fmod(ra_d + kD2PI, kD2PI);
  if (std::isnan(theta) || std::isnan(phi)) return -1;
  const long v = ring_pixel(nside, theta, phi);
  if (v < 0 || v > 12 * nside * nside - 1) return -1;
  return v;
}

void host_det_const(double lel, double xel, double angl, double *offx, double *offy, double *psidet,
                    double *th_, double *ph_, double *sin_th, double *cos_th, double *drp2) {
  // read_bolo_offsets stores {xel, lel, angl}; 460-462 convert them.
  const double o0 = xel, o1 = lel, o2 = angl;
  *offx = o1 / 180.0 * kD2PI / 2.0;
  *offy = o0 / 180.0 * kD2PI / 2.0;
  *psidet = o2 / 180.0 * kD2PI / 2.0;
  const double x = *offx, y = *offy;
  *th_ = std::asin(std::sqrt(x * x + y * y));
  *ph_ = std::atan2(y, x);
  *sin_th = std::sin(*th_);
  *cos_th = std::cos(*th_);
  *drp2 = -std::sin(*th_) * std::cos(*ph_);
}

void host_spn_row(const double *ispf, long nssp, long ns, long count, double *row) {
  // 533-537, including the truncating index expression; long rows are
  // split over host threads (each entry is independent)
  auto fill = [&](long a, long b) {
    for (long ii = a; ii < b; ii++) row[ii] = ispf[long(double(ii) / nssp * ns)] / double(ns) / double(ns);
  };
  const long nth = count >= (1L << 18) ? 8 : 1;
  if (nth == 1) {
    fill(0, count);
  } else {
    std::vector<std::thread> pool;
    for (long w = 0; w < nth; w++) pool.emplace_back(fill, count * w / nth, count * (w + 1) / nth);
    for (auto &t : pool) t.join();
  }
  if (count > 0) row[0] = ispf[1] * 1e-6 / double(ns) / double(ns);
}

namespace {
int64_t isqrt64(int64_t v) {
  int64_t r = (int64_t)std::sqrt((double)v);
  while (r * r > v) r--;
  while ((r + 1) * (r + 1) <= v) r++;
  return r;
}
}  // namespace

int host_ring2xyf(long nside, long pix, int *face, int *x, int *y) {
  if (nside < 1 || (nside & (nside - 1))) return -1;
  const int64_t ns = nside, npix = 12 * ns * ns, ncap = 2 * ns * (ns - 1);
  if (pix < 0 || pix >= npix) return -1;
  // centre (z, phi) of the ring pixel
  double z, phi;
  if (pix < ncap) {
    const int64_t ir = (1 + isqrt64(1 + 2 * (int64_t)pix)) >> 1;
    const int64_t iphi = pix + 1 - 2 * ir * (ir - 1);
    z = 1.0 - (double)(ir * ir) / (3.0 * (double)(ns * ns));
    phi = ((double)iphi - 0.5) * M_PI / (2.0 * (double)ir);
  } else if (pix < npix - ncap) {
    const int64_t ip = pix - ncap;
    const int64_t ir = ip / (4 * ns) + ns;
    const int64_t iphi = ip % (4 * ns) + 1;
    const double fodd = ((ir + ns) & 1) ? 1.0 : 0.5;
    z = (double)(2 * ns - ir) * 2.0 / (3.0 * (double)ns);
    phi = ((double)iphi - fodd) * M_PI / (2.0 * (double)ns);
  } else {
    const int64_t ip = npix - pix;
    const int64_t ir = (1 + isqrt64(2 * ip - 1)) >> 1;
    const int64_t iphi = 4 * ir + 1 - (ip - 2 * ir * (ir - 1));
    z = -1.0 + (double)(ir * ir) / (3.0 * (double)(ns * ns));
    phi = ((double)iphi - 0.5) * M_PI / (2.0 * (double)ir);
  }
  // face and in-face coordinates of that direction (nested construction)
  const double za = std::fabs(z);
  double tt = std::fmod(phi * (2.0 / M_PI), 4.0);
  if (tt < 0.0) tt += 4.0;
  if (za <= 2.0 / 3.0) {
    const double t1 = (double)ns * (0.5 + tt), t2 = (double)ns * (z * 0.75);
    const int64_t jp = (int64_t)(t1 - t2), jm = (int64_t)(t1 + t2);
    const int64_t ifp = jp / ns, ifm = jm / ns;
    *face = (int)(ifp == ifm ? (ifp | 4) : (ifp < ifm ? ifp : ifm + 8));
    *x = (int)(jm & (ns - 1));
    *y = (int)(ns - (jp & (ns - 1)) - 1);
  } else {
    const int ntt = std::min(3, (int)tt);
    const double tp = tt - ntt;
    const double tmp = (double)ns * std::sqrt(3.0 * (1.0 - za));
    int64_t jp = (int64_t)(tp * tmp), jm = (int64_t)((1.0 - tp) * tmp);
    jp = std::min<int64_t>(jp, ns - 1);
    jm = std::min<int64_t>(jm, ns - 1);
    if (z >= 0) {
      *face = ntt;
      *x = (int)(ns - jm - 1);
      *y = (int)(ns - jp - 1);
    } else {
      *face = ntt + 8;
      *x = (int)jp;
      *y = (int)jm;
    }
  }
  return 0;
}
