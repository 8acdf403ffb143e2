// sanepic_main.cpp -- drop-in `sanepic` executable on top of libsanepic_hip.
//
// Same command line as patanch/SANEPIC's src/sanepic.cc (getopt string at
// 315-317, run_sanepic.sh:55-73), same input files (segment list, pointing,
// TODs, iSpf spectra, bolometer file), same stdout progress lines and the
// same output map files (Appendix of the format: write_reduced_map 2364-2388,
// names at 1618-1801).  The MPI master/worker split of the reference is
// replaced by one process driving one GPU; when launched under mpirun, ranks
// other than 0 exit 0 at once so `mpirun -n N sanepic ...` still works.
#include <getopt.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <list>
#include <string>
#include <vector>

#include "sanepic_hip.h"

namespace {

constexpr double kD2PI = 6.2831853071795864769252867665590057683943387987502;

void usage(const char *name) {
  std::cerr
      << "usage: " << name << " [options]   (GLS I/Q/U map-making on one MI355X)\n\n"
      << "Inputs\n"
      << "  -F dir      directory holding the detector timelines data_<detector><-B>.bi\n"
      << "  -B ext      suffix of the timeline file names\n"
      << "  -Z prefix   path prefix of the boresight files <prefix>{ra,dec,psi}<-P>.bi\n"
      << "  -P ext      suffix of the boresight file names\n"
      << "  -f file     segment list: int64 (first, last) sample pairs\n"
      << "  -n count    number of segments in that list\n"
      << "  -l length   samples per segment (the same for every segment)\n"
      << "  -k prefix   path prefix of the inverse noise spectra <prefix><detector>\n"
      << "  -u count    values read from each inverse noise spectrum\n"
      << "  -X file     detector table: num pnum lel xel angle tel name (degrees)\n"
      << "  -C name     use this detector (repeat for more)\n"
      << "  -D file     use the detectors listed in this file, one per line\n"
      << "  -d count    number of detectors to use\n"
      << "  -o 0|1      subtract the orbital dipole read from <-F>/orbdip_<detector>\n"
      << "  -b 0|1      also solve for a leakage map L that each detector sees times its coefficient\n"
      << "  -K file     the leakage coefficients, one float64 per detector (normalised on reading)\n"
      << "Map and solver\n"
      << "  -N nside    HEALPix resolution of the maps (RING ordering)\n"
      << "  -p 1        solve for I, Q and U (required)\n"
      << "  -h 0|1      half-wave plate in the optical path\n"
      << "  -w omega    half-wave plate angle step per sample (radians)\n"
      << "  -I file     start from this compressed [I|Q|U] map (an earlier _A_ output)\n"
      << "  -r 0|1      alternate map solves with fits of a gain per detector and segment, until a\n"
      << "              solve starts converged (maps _<pass>, gains relCalib.dat<pass> in the working\n"
      << "              directory; SANEPIC_MAX_RLOOP caps the passes)\n"
      << "  -R 0|1      with -r: one gain per detector instead\n"
      << "Outputs\n"
      << "  -O prefix   path prefix of every output map\n"
      << "  -e tag      tag inserted in the output file names\n"
      << "  -i 0|1      also write the I, Q, U maps of intermediate iterations\n"
      << "  -j step     iterations between intermediate maps (default 10)\n"
      << "Read for compatibility and without effect on the maps: -H -J -A -m -G -L -g -z -M -c\n"
      << "Rejected unless 0 (not available): -E\n";
  exit(1);
}

int mpi_rank() {
  const char *vars[] = {"OMPI_COMM_WORLD_RANK", "PMI_RANK", "PMIX_RANK", "MV2_COMM_WORLD_RANK", "SLURM_PROCID"};
  for (const char *v : vars) {
    const char *s = getenv(v);
    if (s && *s) return atoi(s);
  }
  return 0;
}

// ReadDoubleVECT / ReadLongVECT (3087-3132): fseek(imin * 8) then fread.
template <typename T>
bool read_vect(T *data, const std::string &fn, long imin, long imax) {
  FILE *fp = fopen(fn.c_str(), "rb");
  if (!fp) {
    fprintf(stderr, "ERROR: cannot open %s\n", fn.c_str());
    return false;
  }
  const long nn = imax - imin + 1;
  if (fseek(fp, imin * (long)sizeof(T), SEEK_SET) != 0) {
    fclose(fp);
    fprintf(stderr, "ERROR: cannot seek in %s\n", fn.c_str());
    return false;
  }
  const size_t got = fread(data, sizeof(T), nn, fp);
  fclose(fp);
  if ((long)got != nn) {
    fprintf(stderr, "ERROR: %s holds %zu of the %ld values requested from %ld\n", fn.c_str(), got, nn, imin);
    return false;
  }
  return true;
}

// A host buffer of n doubles, page-locked through the library when it can
// be: the host-to-device copies from it then run at the DMA engines' rate
// instead of through the runtime's pageable staging.  Pageable otherwise.
class HostBuf {
 public:
  explicit HostBuf(size_t n) {
    void *q = nullptr;
    if (sanepic_host_alloc(n * sizeof(double), &q) == SANEPIC_OK && q) {
      p_ = static_cast<double *>(q);
      pinned_ = true;
    } else {
      v_.resize(n);
      p_ = v_.data();
    }
  }
  ~HostBuf() {
    if (pinned_) sanepic_host_free(p_);
  }
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
  double *data() { return p_; }

 private:
  double *p_ = nullptr;
  bool pinned_ = false;
  std::vector<double> v_;
};

// read_bolofile (3134-3157): first word of every non-empty, non-# line.
void read_bolofile(const std::string &fname, std::list<std::string> &bolos) {
  std::ifstream in(fname.c_str());
  if (!in.is_open()) {
    std::cerr << "Error opening bolometer file '" << fname << "'. Exiting.\n";
    exit(1);
  }
  std::string line;
  while (std::getline(in, line)) {
    line.erase(0, line.find_first_not_of(" \t"));
    if (line.empty() || line[0] == '#') continue;
    bolos.push_back(line.substr(0, line.find_first_of(" \t")));
  }
}

// read_bolo_offsets (3159-3188): fields num pnum lel xel angl tel name; the
// last line with the name wins.
void read_bolo_offsets(const std::string &bolo, const std::string &file, double &lel, double &xel,
                       double &angl) {
  FILE *fp = fopen(file.c_str(), "r");
  if (!fp) {
    std::cerr << "ERROR: Can't find offset file. Exiting. \n";
    exit(1);
  }
  long num, pnum;
  double a, b2, c;
  char tel[300], name[300];
  bool found = false;
  while (fscanf(fp, "%ld%ld%lf%lf%lf%299s%299s\n", &num, &pnum, &a, &b2, &c, tel, name) == 7) {
    if (bolo == name) {
      found = true;
      lel = a;
      xel = b2;
      angl = c;
    }
  }
  fclose(fp);
  if (!found) {
    std::cerr << "Bolometer name not found in offset list" << std::endl;
    exit(1);
  }
}

// write_reduced_map (2364-2388): full sky, 0 for unseen pixels.
void write_reduced_map(const std::string &fn, long nside, const double *mapr, const std::vector<int64_t> &indpix) {
  std::vector<double> full(12 * nside * nside, 0.0);
  for (size_t i = 0; i < full.size(); i++)
    if (indpix[i] != -1) full[i] = indpix[i] < 0 ? mapr[-indpix[i] - 2] : mapr[indpix[i]];
  FILE *fp = fopen(fn.c_str(), "wb");
  if (!fp || fwrite(full.data(), sizeof(double), full.size(), fp) != full.size()) {
    printf("WARNING, PROBLEM WITH WRITTEN DATA\n");
  }
  if (fp) fclose(fp);
}

void write_raw(const std::string &fn, const double *v, long n) {
  FILE *fp = fopen(fn.c_str(), "wb");
  if (!fp || fwrite(v, sizeof(double), n, fp) != (size_t)n) printf("WARNING, PROBLEM WITH WRITTEN DATA\n");
  if (fp) fclose(fp);
}

std::string fmt(const char *f, const std::string &a, const std::string &b) {
  char buf[4096];
  snprintf(buf, sizeof buf, f, a.c_str(), b.c_str());
  return buf;
}
std::string fmt(const char *f, const std::string &a, const std::string &b, int k) {
  char buf[4096];
  snprintf(buf, sizeof buf, f, a.c_str(), b.c_str(), k);
  return buf;
}

void die_lib(sanepic_ctx *ctx, int rc, const char *what) {
  fprintf(stderr, "sanepic: %s failed (%d): %s\n", what, rc, sanepic_last_error(ctx));
  exit(1);
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char *argv[]) {
  const int rank = mpi_rank();
  int iterw = 10;
  bool polar = false, hwp = false, initMapOn = false, writeIterOn = true, pfile = false, odip = false;
  int recalib = 0, recalibGlobal = 0, galbp = 0;
  long ndet = 0, nsegtot = 0, ns = 0, nssp = 0, nside = 0;
  double omega = 0.0;
  std::string outdir, bext, pext, termin_out, noisepre, filesegnum, dirfile, dirfilep, initMap, filedetinfo,
      relfgal_file;
  std::list<std::string> bolos;
  int opt;
  while ((opt = getopt(argc, argv,
                       "F:Z:f:n:l:d:D:C:H:h:w:J:o:O:B:R:X:G:P:e:p:A:m:k:K:u:"
                       "c:N:n:L:g:r:M:E:j:z:b:I:i:")) != -1) {
    switch (opt) {
      case 'F': dirfile = optarg; break;
      case 'Z': dirfilep = optarg; pfile = true; break;
      case 'f': filesegnum = optarg; break;
      case 'p': polar = atoi(optarg); break;
      case 'd': ndet = atoi(optarg); break;
      case 'D': read_bolofile(optarg, bolos); std::cerr << "num bolos: " << bolos.size() << std::endl; break;
      case 'C': bolos.push_back(optarg); break;
      case 'O': outdir = optarg; break;
      case 'B': bext = optarg; break;
      case 'X': filedetinfo = optarg; break;
      case 'P': pext = optarg; break;
      case 'e': termin_out = optarg; break;
      case 'k': noisepre = optarg; break;
      case 'u': nssp = atoi(optarg); break;
      case 'N': nside = atoi(optarg); break;
      case 'n': nsegtot = atoi(optarg); break;
      case 'l': ns = atoi(optarg); break;
      case 'j': iterw = atoi(optarg); break;
      case 'h': hwp = atoi(optarg); break;
      case 'w': omega = atof(optarg); break;
      case 'I': initMap = optarg; initMapOn = true; break;
      case 'i': writeIterOn = atoi(optarg); break;
      case 'E':
        if (atoi(optarg)) {  // 513-515
          printf("NOT IMPLEMENTED YET\n");
          exit(0);
        }
        break;
      case 'o': odip = atoi(optarg) != 0; break;
      case 'r': recalib = atoi(optarg); break;
      case 'b': galbp = atoi(optarg); break;
      // -K names the -b file; the rest are read by the reference and never used
      case 'K': relfgal_file = optarg; break;
      case 'R': recalibGlobal = atoi(optarg); break;
      case 'H': case 'J': case 'A': case 'm': case 'G': case 'L': case 'g': case 'z': case 'M': case 'c':
        break;
      default:
        std::cerr << "Option '" << (char)opt << "' not valid. Exiting.\n\n";
        usage(argv[0]);
    }
  }
  if (rank != 0) return 0;  // one process drives the GPU
  if (!pfile) dirfilep = dirfile;
  std::vector<std::string> bolonames(bolos.begin(), bolos.end());
  if ((long)bolonames.size() < ndet || ndet < 1) {
    fprintf(stderr, "sanepic: -d %ld but %zu detectors given\n", ndet, bolonames.size());
    exit(1);
  }
  bolonames.resize(ndet);
  if (!polar) {
    fprintf(stderr, "sanepic: only -p 1 is supported (the reference reads psi unconditionally)\n");
    exit(1);
  }
  if (recalib && galbp) {  // 1122-1123, 1237-1239, 1567-1569 (forceLeakzero, 3262-3275)
    fprintf(stderr, "sanepic: -r with -b (leakage forced to zero at high latitude) is not supported by this build\n");
    exit(1);
  }
  const double t0 = now();
  // detector offsets (458-464)
  std::vector<double> lel(ndet), xel(ndet), angl(ndet);
  for (long i = 0; i < ndet; i++) {
    read_bolo_offsets(bolonames[i], filedetinfo, lel[i], xel[i], angl[i]);
    printf("Angle detector %ld = %10.15g\n", i, angl[i] / 180.0 * kD2PI / 2.0);
  }
  printf("nsegtot = %ld\n", nsegtot);
  // segment list (494-502) and uniform length check (668-672)
  std::vector<long> seg(2 * nsegtot);
  if (!read_vect(seg.data(), filesegnum, 0, 2 * nsegtot - 1)) exit(1);
  for (long i = 0; i < nsegtot; i++)
    if (seg[2 * i + 1] - seg[2 * i] + 1 != ns) {
      printf("Problem with segment size: %ld, %ld\n", seg[2 * i + 1] - seg[2 * i], ns);
      exit(1);
    }
  sanepic_config cfg{};
  cfg.nside = nside;
  cfg.ndet = ndet;
  cfg.nseg = nsegtot;
  cfg.ns = ns;
  cfg.polar = 1;
  cfg.hwp = hwp ? 1 : 0;
  cfg.omega_hwp = omega;
  cfg.precision = SANEPIC_FP64;
  if (const char *p = getenv("SANEPIC_PRECISION")) cfg.precision = strcmp(p, "mixed") == 0 ? SANEPIC_MIXED : SANEPIC_FP64;
  cfg.device = 0;
  sanepic_ctx *ctx = nullptr;
  int rc = sanepic_create(&cfg, &ctx);
  if (rc) die_lib(nullptr, rc, "sanepic_create");
  const int nc = galbp ? 4 : 3;
  if (galbp) {
    // leakage coefficients: ndet doubles from the -K file, normalised to zero
    // mean and unit (population) variance over the detectors (551-575).  The
    // reference's ReadVECT (3061-3085) reads into a buffer of its own and
    // leaves relfGal at 0, so its maps turn NaN; here the file is read.
    std::vector<double> relf(ndet);
    if (relfgal_file.empty() || !read_vect(relf.data(), relfgal_file, 0, ndet - 1)) {
      fprintf(stderr, "sanepic: -b 1 needs -K <file of %ld leakage coefficients>\n", ndet);
      exit(1);
    }
    double ttmean = 0.0, ttsig = 0.0;
    for (long i = 0; i < ndet; i++) ttmean += relf[i] / double(ndet);
    for (long i = 0; i < ndet; i++) ttsig += (relf[i] - ttmean) * (relf[i] - ttmean) / double(ndet);
    for (long i = 0; i < ndet; i++) relf[i] = (relf[i] - ttmean) / sqrt(ttsig);
    for (long i = 0; i < ndet; i++) printf("!!!!!!!!!!!!!!!!!!!!!!!!!!!!! relfGal[%ld] = %10.15g\n", i, relf[i]);
    rc = sanepic_set_leakage(ctx, relf.data());
    if (rc) die_lib(ctx, rc, "sanepic_set_leakage");
  }
  // noise spectra (521-538)
  std::vector<double> spec(nssp);
  for (long i = 0; i < ndet; i++) {
    const std::string fn = noisepre + bolonames[i];
    printf("NoiseSpFiledet = %s\n", fn.c_str());
    if (!read_vect(spec.data(), fn, 0, nssp - 1)) exit(1);
    rc = sanepic_set_detector(ctx, i, lel[i], xel[i], angl[i], spec.data(), nssp);
    if (rc) die_lib(ctx, rc, "sanepic_set_detector");
  }
  printf("NO CORRELATIONS BETWEEN DETECTORS INCLUDED\n");
  printf("%ld\n", nside);
  // pointing and TODs (667-817)
  double tio = 0.0;
  HostBuf ra(ns), dec(ns), psi(ns), tod(ns);
  for (long s = 0; s < nsegtot; s++) {
    const double ta = now();
    const std::string base = dirfilep;
    if (!read_vect(ra.data(), base + "ra" + pext + ".bi", seg[2 * s], seg[2 * s + 1]) ||
        !read_vect(dec.data(), base + "dec" + pext + ".bi", seg[2 * s], seg[2 * s + 1]) ||
        !read_vect(psi.data(), base + "psi" + pext + ".bi", seg[2 * s], seg[2 * s + 1]))
      exit(1);
    tio += now() - ta;
    rc = sanepic_set_pointing(ctx, s, ra.data(), dec.data(), psi.data());
    if (rc) die_lib(ctx, rc, "sanepic_set_pointing");
    for (long i = 0; i < ndet; i++) {
      const double tb = now();
      const std::string fn = dirfile + "/data_" + bolonames[i] + bext + ".bi";
      if (s == 0) printf("%s\n", fn.c_str());
      if (!read_vect(tod.data(), fn, seg[2 * s], seg[2 * s + 1])) exit(1);
      tio += now() - tb;
      rc = sanepic_set_tod(ctx, i, s, tod.data());
      if (rc) die_lib(ctx, rc, "sanepic_set_tod");
      if (odip) {  // 702-720: <data dir>/orbdip_<bolo>, no extension
        const double tc = now();
        if (!read_vect(tod.data(), dirfile + "/orbdip_" + bolonames[i], seg[2 * s], seg[2 * s + 1])) exit(1);
        tio += now() - tc;
        rc = sanepic_set_dipole(ctx, i, s, tod.data());
        if (rc) die_lib(ctx, rc, "sanepic_set_dipole");
      }
    }
  }
  printf("DATA SUCCESSFULLY READ\n");
  int64_t npix = 0, ndef = 0;
  rc = sanepic_pixelize(ctx, &npix, &ndef);
  if (rc) die_lib(ctx, rc, "sanepic_pixelize");
  const long npixtot = nc * npix;
  printf("TOTAL NUMBER OF DETECTORS: %d\n", (int)ndet);
  printf("TOTAL NUMBER OF DATA SEGMENTS: %d\n", (int)nsegtot);
  printf("NUMBER OF SAMPLE PER DATA SEGMENT: %d\n", (int)ns);
  printf("TOTAL NUMBER OF FILLED PIXELS IN ONE MAP: %ld, TOTAL NUMBER: %ld, FOR PROCESSOR %d: %ld, NB OF HEALPIX "
         "PIXEL: %ld \n",
         (long)npix, npixtot, 0, npixtot, 12 * nside * nside);
  printf("Using input noise \n");
  std::vector<int64_t> indpix(12 * nside * nside);
  sanepic_get_indpix(ctx, indpix.data());
  if (recalib) {  // the TOD and the dipole stay on the device for the gain fits
    rc = sanepic_set_recalib(ctx, 1);
    if (rc) die_lib(ctx, rc, "sanepic_set_recalib");
  }
  rc = sanepic_setup(ctx);
  if (rc) die_lib(ctx, rc, "sanepic_setup");
  printf("PtNd Computed...\n");
  std::vector<double> x0;
  if (initMapOn) {  // 1209-1227: compressed npixtot vector
    x0.resize(npixtot);
    if (!read_vect(x0.data(), initMap, 0, npixtot - 1)) exit(1);
  }
  sanepic_solve_opts so{};
  so.precond = SANEPIC_PRECOND_DIAG;
  sanepic_unet *unet = nullptr;
  if (const char *p = getenv("SANEPIC_PRECOND")) {
    if (strcmp(p, "block") == 0) so.precond = SANEPIC_PRECOND_BLOCK3;
    if (strcmp(p, "learned") == 0) {
      // trained parameters (float32, the library's layout), e.g. written by
      // python -m sanepic_amd.train_precond for this scan
      const char *w = getenv("SANEPIC_UNET_WEIGHTS");
      if (!w) {
        fprintf(stderr, "SANEPIC_PRECOND=learned needs SANEPIC_UNET_WEIGHTS=<parameter file>\n");
        exit(1);
      }
      // the network's own nside (a power of two dividing the map's; default
      // min(nside, 64), as python -m sanepic_amd.train_precond trains it)
      long unside = nside < 64 ? nside : 64;
      if (const char *e = getenv("SANEPIC_UNET_NSIDE")) unside = atol(e);
      rc = sanepic_unet_create(unside, cfg.device, &unet);
      if (rc) {
        fprintf(stderr, "sanepic_unet_create: %s\n", sanepic_unet_last_error(nullptr));
        exit(1);
      }
      int64_t np = 0;
      sanepic_unet_nparams(unet, &np);
      std::vector<float> par(np);
      FILE *fw = fopen(w, "rb");
      if (!fw || fread(par.data(), sizeof(float), np, fw) != (size_t)np) {
        fprintf(stderr, "cannot read %ld parameters from %s\n", (long)np, w);
        exit(1);
      }
      fclose(fw);
      if (sanepic_unet_set_params(unet, par.data()) || sanepic_learn_attach(ctx, unet)) {
        fprintf(stderr, "learned preconditioner: %s\n", sanepic_last_error(ctx));
        exit(1);
      }
      so.precond = SANEPIC_PRECOND_LEARNED;
      so.deflate = 1;
    }
  }
  so.refresh = 10;
  // Q/U entries of pixels with fewer than 4 polarised hits are out of every
  // [3 npix] vector (1316-1330): |b|^2 below is taken over the rest
  std::vector<int64_t> hits(2 * npix);
  sanepic_get_setup(ctx, hits.data(), nullptr, nullptr);
  auto masked_norm2 = [&](const std::vector<double> &v) {
    double n2 = 0.0;
    for (long i = 0; i < npixtot; i++) {
      const long p = i % npix, c = i / npix;
      if ((c == 1 || c == 2) && hits[npix + p] < 4) continue;
      n2 += v[i] * v[i];
    }
    return n2;
  };
  // The recalibration loop (1041-2316): each pass solves the map with the
  // current relCalib, writes it with the pass number, and with -r fits the
  // gains of every (det, seg) block to it.  The loop ends when a pass starts
  // from a map that already solves its system to 1e-15 (var0 / var00, 1042),
  // after 499 passes, or at SANEPIC_MAX_RLOOP passes.
  const long nblk = ndet * nsegtot;
  std::vector<double> gains(nblk, 1.0);  // [det][seg]
  int rloop_cap = 500;
  if (const char *e = getenv("SANEPIC_MAX_RLOOP")) rloop_cap = std::min(500, atoi(e) + 1);
  const double StopCriterium = 1e-15;
  double var0 = 1.0, var00 = 1.0, var00_first = 0.0, tsetup = 0.0, tloop = 0.0;
  int iter_total = 0;
  std::vector<double> S(npixtot), aux(npixtot), bvec(npixtot);
  int rloop = 1;
  while (rloop && rloop < rloop_cap && var0 / var00 > StopCriterium) {
    if (rloop > 1) {
      rc = sanepic_set_gains(ctx, gains.data());
      if (rc) die_lib(ctx, rc, "sanepic_set_gains");
    }
    // the first pass keeps its own |b|^2 as the denominator of crit (1377-1385)
    so.var00 = rloop > 1 ? var00_first : 0.0;
    const double *xstart = rloop > 1 ? S.data() : (initMapOn ? x0.data() : nullptr);
    rc = sanepic_solve_begin(ctx, &so, xstart);
    if (rc == SANEPIC_EDATA) {
      printf("%s\n", sanepic_last_error(ctx));
      exit(0);  // the reference exits 0 here (1344-1348)
    }
    if (rc) die_lib(ctx, rc, "sanepic_solve_begin");
    // var0 = |b - A S|^2 of the starting map, var00 = |b|^2 (1360-1399)
    sanepic_get_rhs(ctx, bvec.data());
    const double bb = masked_norm2(bvec);
    if (rloop == 1) var00 = var00_first = bb;
    var0 = bb;
    if (xstart) {
      double rel = 0.0;
      rc = sanepic_true_residual(ctx, &rel);
      if (rc) die_lib(ctx, rc, "sanepic_true_residual");
      var0 = rel * rel * bb;
    }
    printf("var0 = %lg\n", var0);
    printf("var00 = %lg\n", var00);
    if (rloop == 1) tsetup = now() - t0;
    // 1415-1761
    int iter = 0;
    double crit = 1.0;
    const double tl0 = now();
    while ((iter < 2000 && crit > StopCriterium) || iter == 0) {
      sanepic_iter_info info{};
      rc = sanepic_solve_step(ctx, &info);
      if (rc) die_lib(ctx, rc, "sanepic_solve_step");
      crit = info.crit;
      printf("iter = %d, crit = %10.15g, crit2 = %10.15g, chi2 = %10.15g\n", iter, info.crit, info.crit2, info.chi2);
      fflush(stdout);
      if (iter == 1 && rloop == 1) {  // 1616-1707
        std::vector<double> tmp(npix), sumc(npixtot);
        sanepic_get_nvar(ctx, aux.data());
        write_reduced_map(fmt("%s_NVAR%s", outdir, termin_out), nside, aux.data(), indpix);
        sanepic_get_sumcross(ctx, sumc.data());
        for (long i = 0; i < npix; i++) tmp[i] = double(hits[i]);
        write_reduced_map(fmt("%s_HITS%s", outdir, termin_out), nside, tmp.data(), indpix);
        write_reduced_map(fmt("%s_COS2P_%s", outdir, termin_out), nside, sumc.data(), indpix);
        write_reduced_map(fmt("%s_SIN2P_%s", outdir, termin_out), nside, sumc.data() + npix, indpix);
        write_reduced_map(fmt("%s_Q_NVAR%s", outdir, termin_out), nside, aux.data() + npix, indpix);
        write_reduced_map(fmt("%s_U_NVAR%s", outdir, termin_out), nside, aux.data() + 2 * npix, indpix);
        if (galbp)  // 1693-1697
          write_reduced_map(fmt("%s_Leak_NVAR%s", outdir, termin_out), nside, aux.data() + 3 * npix, indpix);
        for (long i = 0; i < npix; i++) tmp[i] = double(hits[i + npix]);
        write_reduced_map(fmt("%s_POL_HITS%s", outdir, termin_out), nside, tmp.data(), indpix);
      }
      if (iterw && (iter % iterw) == 0 && writeIterOn && rloop == 1) {  // 1711-1751
        sanepic_get_map(ctx, S.data());
        write_reduced_map(fmt("%s_I%s_%d", outdir, termin_out, iter), nside, S.data(), indpix);
        write_reduced_map(fmt("%s_Q%s_%d", outdir, termin_out, iter), nside, S.data() + npix, indpix);
        write_reduced_map(fmt("%s_U%s_%d", outdir, termin_out, iter), nside, S.data() + 2 * npix, indpix);
        if (galbp)  // 1738-1742
          write_reduced_map(fmt("%s_Leak%s_%d", outdir, termin_out, iter), nside, S.data() + 3 * npix, indpix);
      }
      iter++;
    }
    printf("\n");
    tloop += now() - tl0;
    iter_total += iter;
    // 1773-1817
    sanepic_get_map(ctx, S.data());
    write_raw(fmt("%s_A_%s_%d", outdir, termin_out, rloop), S.data(), npixtot);
    write_reduced_map(fmt("%s_I_%s_%d", outdir, termin_out, rloop), nside, S.data(), indpix);
    write_reduced_map(fmt("%s_Q_%s_%d", outdir, termin_out, rloop), nside, S.data() + npix, indpix);
    write_reduced_map(fmt("%s_U_%s_%d", outdir, termin_out, rloop), nside, S.data() + 2 * npix, indpix);
    if (galbp)  // 1803-1806
      write_reduced_map(fmt("%s_Leak_%s_%d", outdir, termin_out, rloop), nside, S.data() + 3 * npix, indpix);
    printf("Map-making Ends \n");
    printf("JUST BEFORE recalib \n");
    if (recalib) {
      // gain fit (2157-2316): per block zeta / gamma, the minimiser of
      // MPI_compute_chi2; with -R 1 one gain per detector from the sums over
      // its blocks.  The reference's own ratio carries factors 2 and 1/2 in
      // zeta and gamma and a map scale absrecal read from values only -R
      // sets; with those equal and absrecal = 1 its 50 refits of a pass are
      // identical, so one is made.
      std::vector<double> sums(3 * nblk), g2(nblk);
      rc = sanepic_recal_sums(ctx, S.data(), sums.data());
      if (rc) die_lib(ctx, rc, "sanepic_recal_sums");
      auto chi2_of = [&](const std::vector<double> &g) {  // 3206-3255
        double c2 = 0.0;
        for (long is = 0; is < nsegtot; is++)
          for (long id = 0; id < ndet; id++) {
            const long b = id * nsegtot + is;
            const double *sm = sums.data() + 3 * b;
            c2 += 2.0 * (sm[2] - 2.0 * g[b] * sm[0] + g[b] * g[b] * sm[1]);
          }
        return c2;
      };
      printf("Before recal chi2 = %10.15g\n", chi2_of(gains));
      for (long id = 0; id < ndet; id++) {
        if (recalibGlobal) {
          double za = 0.0, ga = 0.0;
          for (long is = 0; is < nsegtot; is++) {
            za += sums[3 * (id * nsegtot + is)];
            ga += sums[3 * (id * nsegtot + is) + 1];
          }
          for (long is = 0; is < nsegtot; is++) g2[id * nsegtot + is] = za / ga;
          printf("Coef calib det %ld is %10.15g\n", id, za / ga);
        } else {
          for (long is = 0; is < nsegtot; is++) {
            const long b = id * nsegtot + is;
            g2[b] = sums[3 * b] / sums[3 * b + 1];
          }
        }
      }
      printf("After recal chi2 = %10.15g\n", chi2_of(g2));
      gains = g2;
      // relCalib.dat<rloop> in the working directory: relCalib[idet + iseg ndet]
      std::vector<double> rel(nblk);
      for (long id = 0; id < ndet; id++)
        for (long is = 0; is < nsegtot; is++) rel[id + is * ndet] = gains[id * nsegtot + is];
      write_raw("relCalib.dat" + std::to_string(rloop), rel.data(), nblk);
    }
    rloop += 1;
    if (!recalib) rloop = 0;
  }
  fprintf(stderr,
          "sanepic: %d iterations, set-up %.3f s (file I/O %.3f s), CG loop %.3f s (%.3f ms/iter), "
          "%ld deferred pixel samples, %.2f GB on device\n",
          iter_total, tsetup, tio, tloop, 1e3 * tloop / std::max(iter_total, 1), (long)ndef,
          [&] { int64_t b = 0; sanepic_device_bytes(ctx, &b); return b / 1e9; }());
  sanepic_destroy(ctx);
  sanepic_unet_destroy(unet);
  return 0;
}
