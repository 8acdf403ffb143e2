// pixel_host.h -- host-side pieces of the set-up that must use the platform
// libm bit for bit (see pixel_host.cpp).
#pragma once

// RING pixel of one sample from its boresight (ra, dec, psi) and the
// detector offsets in radians; -1 when the angles are NaN or the pixel is out
// of range (the reference then keeps the previous pixel, 862-872).
long host_sample_pixel(long nside, double offx, double offy, double ra, double dec, double psi);

// Detector constants of 458-463 and pointingshift (3419, 3435).
void host_det_const(double lel, double xel, double angl, double *offx, double *offy, double *psidet,
                    double *th_, double *ph_, double *sin_th, double *cos_th, double *drp2);

// SpN_all diagonal row of one detector (533-537): its first `count` entries
// (the weights are even, W[k] = SpN[min(k, ns - k)], so count = ns/2 + 1
// serves the whole spectrum).
void host_spn_row(const double *ispf, long nssp, long ns, long count, double *row);

// Nested base-face coordinates of a RING pixel (nside a power of two): face
// f in 0..11 and the pixel's column x and row y inside the face, from the
// pixel centre (Gorski et al. 2005, eqs. 4-9 for the centre, and the
// face/x/y construction of the NESTED scheme).  Returns 0, or -1 on bad input.
int host_ring2xyf(long nside, long pix, int *face, int *x, int *y);
