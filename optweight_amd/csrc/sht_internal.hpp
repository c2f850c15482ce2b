// Internal structures shared by the SHT plan, the Legendre kernels and the
// ring-FFT kernels.
#pragma once

#include "common.hpp"

#include <vector>

namespace owmi {

// Exponent step of the scaled Legendre recurrence used at plan creation to
// locate the first non-negligible multipole: a value is stored as
// v * 2^(kScale * k), k <= 0.
template <typename T> struct ScaleStep;
template <> struct ScaleStep<double> { static constexpr int bits = 128; };
template <> struct ScaleStep<float> { static constexpr int bits = 32; };

// Functions below this magnitude are dropped: |slam_lm| < 1e-30 for every
// l below the first significant multipole l_f(m, ring) (the functions grow
// monotonically in l there), far below float64 rounding of O(1) terms.
constexpr double kNegligible = 1e-30;

// Device tables of one precision.
//
// Recurrences are stored normalised: slam_l = c_l mu_l with
//   mu_l = (A'_l x + B'_l) mu_{l-1} - mu_{l-2},
// c_l = C_l c_{l-2} (c = 1 at the first two multipoles). Each (m, ring pair,
// function) starts at its first significant multipole l_f with the seeds
// mu_{l_f - 1}, mu_{l_f}; l_f = lmax + 1 means negligible throughout.
template <typename T>
struct PlanTables {
  T *x = nullptr;          // cos(theta_north) per pair [npairs]
  int *lf0 = nullptr;      // spin-0 l_f per (m, pair) [(mmax+1) * npairs]
  T *s0 = nullptr;         // spin-0 (mu_{lf-1}, mu_lf) per (m, pair)
  int *lfp = nullptr;      // spin +2
  T *sp = nullptr;
  int *lfm = nullptr;      // spin -2
  T *sm = nullptr;
  T *coef0 = nullptr;      // spin-0 (A', c) per alm index [2 * nelem]
  T *coef2 = nullptr;      // spin-2 (A', B', c) per alm index [3 * nelem]
  T *twiddle = nullptr;    // exp(-2 pi i k / Mmax), k < Mmax, complex [Mmax]
  T *chirp = nullptr;      // per distinct ring length: exp(-i pi n^2 / N), complex
  T *bhat = nullptr;       // per distinct ring length: FFT_M(conj chirp, wrapped) / M
  T *ringw = nullptr;      // ring weight per ring (rounded to T)
  T *phi0 = nullptr;       // phi0 per ring
  bool ready = false;
};

// Layouts of the ring coefficients between the Legendre and the FFT stages,
// row = transform * ncomp + component, side 0 / 1 the north / south ring of
// a ring pair:
//  * synthesis (the Legendre synthesis writes, the FFTs read), pair-major:
//      element ((row * (mmax + 1) + m) * npairs + pair) * 2 + side.
//    A Legendre wave (lanes over pairs, one m) writes whole lines; an FFT
//    workgroup (one pair, threads over m) reads both rings of its pair with
//    one access per m.
//  * adjoint (the FFTs write, the Legendre adjoint reads), m-blocked
//    pair-major with blocks of MB = coef_mb<T>() multipoles:
//      element (((row * nmb + m / MB) * npairs + pair) * MB + m % MB) * 2 + side,
//    nmb = ceil((mmax + 1) / MB): the MB values of a pair fill one 64-byte
//    segment, which an FFT workgroup writes whole; an adjoint wave reads both
//    rings of a pair with one access per lane, two lanes per 128-byte line.
//    (Whole-line blocks make the adjoint's reads twice as scattered; no m
//    blocks make the FFT's writes 16-byte pieces of lines; both measured
//    slower. The synthesis needs no blocks: its FFT reads of 16 bytes per
//    line are served by L2, as the workgroups of one XCD take neighbouring
//    pairs.)
template <typename T> constexpr int coef_mb() { return 64 / (4 * (int)sizeof(T)); }
template <typename T>
__host__ __device__ inline long long coef_nmb(int mmax) {
  return (mmax + coef_mb<T>()) / coef_mb<T>();
}
// element index of (row, m, pair, side 0) in the adjoint layout
template <typename T>
__host__ __device__ inline long long coef_pair_at(long long row, int m, long long pair, int mmax,
                                                  long long npairs) {
  constexpr int MB = coef_mb<T>();
  return (((row * coef_nmb<T>(mmax) + m / MB) * npairs + pair) * MB + m % MB) * 2;
}
// element index of (row, m, pair, side 0) in the synthesis layout
__host__ __device__ inline long long coef_syn_at(long long row, int m, long long pair, int mmax,
                                                 long long npairs) {
  return ((row * (mmax + 1) + m) * npairs + pair) * 2;
}

struct owmi_plan_impl {
  int device = 0;
  int lmax = 0, mmax = 0;
  long long nrings = 0, npix = 0, npairs = 0;
  int log2_mmax_fft = 0;  // log2 of the largest Bluestein length
  int mmax_fft = 0;       // largest Bluestein length M
  int uniform_logm = 0;   // log2 of the Bluestein length when all rings share it, else 0

  // host copies
  std::vector<long long> h_nphi, h_offsets;
  std::vector<double> h_theta, h_weight, h_phi0;
  std::vector<int> h_pair_n, h_pair_s;

  // device geometry (precision independent)
  int *pair_n = nullptr;         // north ring of each pair
  int *pair_s = nullptr;         // south ring or -1
  long long *nphi = nullptr;     // per ring
  long long *offsets = nullptr;  // per ring
  int *ring_logm = nullptr;      // log2 of the Bluestein length per ring
  long long *ring_chirp_off = nullptr;  // offset (complex elements) into chirp table
  long long *ring_bhat_off = nullptr;   // offset (complex elements) into bhat table
  // FFT jobs: pairs of rings with equal nphi transformed together
  long long nfft_jobs = 0;
  int *fft_ring_a = nullptr;
  int *fft_ring_b = nullptr;  // -1 when the job holds one ring
  int *ring_slot = nullptr;   // 2 pair + side of each ring (side 0: north, 1: south)

  PlanTables<double> t64;
  PlanTables<float> t32;

  // workspace for ring Fourier coefficients (grows)
  void *work = nullptr;
  size_t work_bytes = 0;
  void *ensure_work(size_t bytes);
  // workspace for per-group partial sums of the adjoint Legendre transform
  void *work2 = nullptr;
  size_t work2_bytes = 0;
  void *ensure_work2(size_t bytes);
  // FFT buffers of rings whose Bluestein length exceeds LDS (grows)
  void *work3 = nullptr;
  size_t work3_bytes = 0;
  void *ensure_work3(size_t bytes);
  // pair-major ring coefficients of the sandwich's analysis half (the
  // synthesis half reads the ring-major ones from work)
  void *work4 = nullptr;
  size_t work4_bytes = 0;
  void *ensure_work4(size_t bytes);
  // Workspaces outgrown by a larger call stay allocated until the plan is
  // destroyed: kernels recorded into a HIP graph by an earlier call keep
  // their addresses.
  std::vector<void *> retired;

  size_t nchirp = 0, nbhat = 0;  // complex element counts of the tables

  // Launch choices set by an autotuner (owmi_sht_plan_tune): an entry
  // applies to transforms of at least min_batch right-hand sides (the entry
  // with the largest such min_batch wins); 0 keeps the library's own rule.
  struct TuneEntry {
    long long min_batch = 1;
    int v[OWMI_TUNE_MFMA64_MIN + 1] = {};  // by OWMI_TUNE_* parameter (1 ...)
  };
  std::vector<TuneEntry> tune;  // sorted by min_batch
  // value of a parameter for a batch of ntrans (0: library rule)
  int tuned(int param, long long ntrans) const {
    int out = 0;
    for (const TuneEntry &e : tune)
      if (e.min_batch <= ntrans) out = e.v[param];
    return out;
  }
};

// Plan-time search of the first significant multipole (legendre.hip):
// from seeds v * 2^(128 k) at l = max(m, |s|) and unnormalised recurrence
// coefficients (spin 0: A, C; spin 2: A, B, C per alm index), writes l_f and
// the float64 values slam_{l_f - 1}, slam_{l_f} per (m, pair) and function.
void launch_seed_finder(owmi_plan_impl *p, const double *xs, const double *v0, const int *k0,
                        const double *vp, const int *kp, const double *vm, const int *km,
                        const double *U0, const double *U2, int *lf0, double *s0, int *lfp,
                        double *sp, int *lfm, double *sm);

// Launchers (legendre.hip)
template <typename T>
void launch_synthesis(owmi_plan_impl *p, int spin, const cplx_t<T> *alm, long long alm_tstride,
                      cplx_t<T> *G, long long ntrans, hipStream_t st);
template <typename T>
void launch_adjoint(owmi_plan_impl *p, int spin, const cplx_t<T> *G, cplx_t<T> *alm,
                    long long alm_tstride, long long ntrans, hipStream_t st);

// Launchers (ringfft.hip). G rows are (transform * ncomp + comp) * nrings + ring,
// each of mmax + 1 complex coefficients.
template <typename T>
void launch_ring_synth(owmi_plan_impl *p, const cplx_t<T> *G, T *map, long long map_tstride,
                       int ncomp, long long ntrans, int weighted, hipStream_t st);
template <typename T>
void launch_ring_adjoint(owmi_plan_impl *p, const T *map, long long map_tstride, cplx_t<T> *G,
                         int ncomp, long long ntrans, int weighted, hipStream_t st);
template <typename T>
void launch_ring_sandwich(owmi_plan_impl *p, const cplx_t<T> *G, cplx_t<T> *Gout,
                          const T *diag, long long diag_tstride, int ncomp, long long ntrans,
                          hipStream_t st);
// Spectral sandwich (ringfft.hip): log2 of its FFT length M (the smallest
// power of two >= 4 mmax + 1, at least 16), or 0 when the plan's FFT tables
// do not cover it; the multipliers [comp][ring][M] of a pixel diagonal
// (comp, npix) for one precision, and the sandwich of the ring
// coefficients with them (synthesis layout in, adjoint layout out).
int conv_logm(const owmi_plan_impl *p);
template <typename T>
void launch_conv_mult(owmi_plan_impl *p, const T *diag, long long diag_cs, int ncomp, T *mult,
                      long long mult_cs, hipStream_t st);
template <typename T>
void launch_ring_conv(owmi_plan_impl *p, const cplx_t<T> *G, cplx_t<T> *Gout, const T *mult,
                      long long mult_cs, int ncomp, long long ntrans, hipStream_t st);
// Fill bhat for every distinct ring length from the chirp table (plan setup).
void launch_bluestein_init(owmi_plan_impl *p, const std::vector<long long> &lens,
                           const std::vector<int> &logms, const std::vector<long long> &chirp_off,
                           const std::vector<long long> &bhat_off);

}  // namespace owmi

struct owmi_sht_plan {
  owmi::owmi_plan_impl impl;
};
