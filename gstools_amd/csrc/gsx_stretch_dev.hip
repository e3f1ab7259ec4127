// Ensemble MCMC of spectral radii on the GPU: one wavefront per chain.
//
// GSTools samples the radii of its spectral modes with emcee's stretch move
// (src/gstools/random/rng.py:38-104, called from generator.py:380-385).  The
// host restatement (gsx_stretch.cpp) draws every random number of a step
// from the chain's legacy numpy MT19937 stream and advances the chains in
// lockstep with one log-probability evaluation per half-move.  This file
// runs the whole chain on the device instead: block r = one 64-lane wave =
// chain r, lane w = walker w (nwalkers <= 64), for every step of burn-in and
// production, with the same draws in the same order:
//
//   per step:  one random_sample (the move choice, discarded);
//              shuffle of the labels arange(nw) % 2 (Fisher-Yates from the
//              last position down, partner j in [0, i] by the masked draw);
//   per half-move (label 0, then 1; ns walkers carry it, nc do not):
//              ns random_sample stretch variates, ns masked draws in
//              [0, nc - 1] for the partners, ns random_sample for the
//              acceptance test;
//              zz = ((a - 1) u + 1)^2 / a,  q = c_p - (c_p - x) zz,
//              accept iff (0 + lp(q)) - lp(x) > log(u_acc).
//
// The MT19937 key (624 words) of the chain sits in LDS.  The twist runs in
// three dependency levels: word i reads words i, i + 1 (old, except word 0
// for i = 623) and (i + 397) mod 624 (old for i < 227, new for i >= 227), so
// [0, 227) needs only the old key, [227, 454) the first level and [454, 624)
// the second: exactly the sequential recurrence.  Draws read the key at the stream position, each
// lane tempering one word; the rejection loops of the masked draws test 64
// candidate words at once and take the first acceptable ones (ballot and
// rank), so the stream advances exactly as the sequential code's does.
// Every arithmetic step of the move is rounded as the host's (no fused
// multiply-add); the log-probability is the model's closed form in the
// order numpy evaluates it, with the device's exp / log / pow (an accept
// decision could only differ where the two sides agree to about an ulp).
#include "gsx_internal.h"

namespace gsx {

namespace {

constexpr int kMtN = 624;
constexpr int kMtM = 397;
constexpr uint32_t kMatrixA = 0x9908b0dfu;
constexpr uint32_t kUpper = 0x80000000u;
constexpr uint32_t kLower = 0x7fffffffu;
constexpr int kWave = 64;
constexpr int kStateWords = kMtN + 1;  // key + position

__device__ __forceinline__ uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

__device__ __forceinline__ uint32_t smear(uint32_t m) {
    m |= m >> 1;
    m |= m >> 2;
    m |= m >> 4;
    m |= m >> 8;
    m |= m >> 16;
    return m;
}

__device__ __forceinline__ int lane_id() { return (int)threadIdx.x; }

// One chain's MT19937 stream, shared by the 64 lanes of its wave.  With one
// wave per SIMD every LDS round trip is exposed, so each draw routine reads
// the key once: lanes load the words they need directly, and the label
// shuffle works on a register window of 64 tempered words.
struct WaveMt {
    uint32_t* key;  // LDS, kMtN words
    int pos;        // stream position (uniform across the wave)
    uint32_t wv;    // window: lane l holds the tempered word wbase + l (l < wlen)
    int wbase, wlen;

    // One dependency level of the twist, words [lo, hi): every lane reads its
    // words into registers, then writes them after a barrier.
    __device__ void twist_level(int lo, int hi) {
        constexpr int kPer = (kMtN - kMtM + kWave - 1) / kWave;  // words per lane and level
        const int l = lane_id();
        uint32_t nv[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = lo + l + k * kWave;
            nv[k] = 0;
            if (i < hi) {
                const uint32_t y = (key[i] & kUpper) | (key[i + 1 < kMtN ? i + 1 : 0] & kLower);
                const int far = i + kMtM < kMtN ? i + kMtM : i + kMtM - kMtN;
                nv[k] = key[far] ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = lo + l + k * kWave;
            if (i < hi) key[i] = nv[k];
        }
        __syncthreads();
    }

    __device__ void twist() {
        constexpr int kLevel = kMtN - kMtM;  // 227
        twist_level(0, kLevel);
        twist_level(kLevel, 2 * kLevel);
        twist_level(2 * kLevel, kMtN);
        pos = 0;
    }

    // advance by n words without reading them
    __device__ void skip(int n) {
        const int avail = kMtN - pos;
        if (n > avail) {
            twist();
            pos = n - avail;
        } else {
            pos += n;
        }
    }

    // random_sample() number i of a batch of `count`: words pos + 2i and
    // pos + 2i + 1, read by the lanes that `need` one; the stream then
    // advances by 2 count words (twisting on the way where the batch crosses
    // the end of the key)
    __device__ double double_at(int i, bool need, int count) {
        const int p0 = pos;
        const int o1 = p0 + 2 * i, o2 = o1 + 1;
        uint32_t a = 0, b = 0;
        if (need) {
            if (o1 < kMtN) a = key[o1];
            if (o2 < kMtN) b = key[o2];
        }
        const int total = 2 * count;
        if (total > kMtN - p0) {
            twist();
            if (need) {
                if (o1 >= kMtN) a = key[o1 - kMtN];
                if (o2 >= kMtN) b = key[o2 - kMtN];
            }
            pos = p0 + total - kMtN;
        } else {
            pos = p0 + total;
        }
        a = temper(a) >> 5;
        b = temper(b) >> 6;
        return ((double)a * 67108864.0 + (double)b) / 9007199254740992.0;
    }

    // register window of the words from pos on (twists at the end of the key)
    __device__ void load_window() {
        if (pos >= kMtN) twist();
        const int l = lane_id();
        wbase = pos;
        wlen = kMtN - pos < kWave ? kMtN - pos : kWave;
        wv = l < wlen ? temper(key[pos + l]) : 0u;
    }

    // Fisher-Yates shuffle of the 1-bit labels of nw walkers: for i = nw-1
    // down to 1, swap labels i and j = masked draw in [0, i].  Calls that
    // share a mask are resolved together: call k of the run tries window
    // word off + k against its own bound; every call before the first
    // rejection is done (its word accepted), the rejecting call then retries
    // from the next word -- the sequential consumption, several calls a pass.
    __device__ void shuffle_labels(unsigned long long& labels, int nw) {
        const int l = lane_id();
        load_window();
        int i = nw - 1;
        while (i >= 1) {
            const int off = pos - wbase;
            if (off >= wlen) {
                load_window();
                continue;
            }
            const uint32_t mask = smear((uint32_t)i);
            const int lo = mask > 1u ? (int)((mask + 1u) >> 1) : 1;  // smallest bound with this mask
            const int k = l - off;
            const bool cand = k >= 0 && k <= i - lo && l < wlen;
            const bool acc = cand && (wv & mask) <= (uint32_t)(i - k);
            const unsigned long long fail = __ballot(cand && !acc);
            const unsigned long long cands = __ballot(cand);
            const int done = fail ? (__ffsll((long long)fail) - 1 - off) : __popcll(cands);
            for (int t = 0; t < done; ++t) {
                const uint32_t j = (uint32_t)__builtin_amdgcn_readlane((int)wv, off + t) & mask;
                const int b = i - t;
                const unsigned long long bi = (labels >> b) & 1ull, bj = (labels >> j) & 1ull;
                if (bi != bj) labels ^= (1ull << b) | (1ull << j);
            }
            i -= done;
            pos += done + (fail ? 1 : 0);
        }
    }

    // `count` masked draws in [0, m] (count <= 64): the draws are left in
    // slot[0 .. count) of LDS (the caller reads them after a barrier); m == 0
    // draws nothing (every value is 0)
    __device__ void masked_batch(uint32_t m, int count, uint32_t* slot) {
        if (m == 0) return;
        const uint32_t mask = smear(m);
        const int l = lane_id();
        const unsigned long long below = (l == 0) ? 0ull : (~0ull >> (64 - l));
        int got = 0;
        load_window();
        for (;;) {
            const uint32_t v = wv & mask;
            const bool ok = l < wlen && v <= m;
            const unsigned long long bal = __ballot(ok);
            const int rank = __popcll(bal & below);
            const int need = count - got;
            if (ok && rank < need) slot[got + rank] = v;
            const int total = __popcll(bal);
            if (total >= need) {
                const unsigned long long last = __ballot(ok && rank == need - 1);
                pos = wbase + __ffsll((long long)last);  // that lane + 1
                return;
            }
            got += total;
            pos = wbase + wlen;
            load_window();
        }
    }
};

}  // namespace

// Log of the radial spectral pdf of a closed-form model at radius |q|:
// log(rad_fac(d, r) |S(r)|), 0 pdf at r <= 1e-8 for d > 1 (numpy's isclose
// to 0), non-finite pdf values set to 0 (covmodel/tools.py:374-405).
__device__ double ln_rad_pdf(const gsx_stretch_model& md, double q) {
#pragma clang fp contract(off)
    const double r = fabs(q);
    if (md.dim > 1 && r <= 1e-8) return -INFINITY;
    double fac;
    if (md.dim == 1) {
        fac = 2.0;
    } else if (md.dim == 2) {
        fac = 6.283185307179586 * r;          // 2 pi r
    } else {
        fac = 12.566370614359172 * (r * r);   // 4 pi r^2
    }
    const double L = md.len_rescaled;
    double s;
    if (md.kind < GSX_STRETCH_GAUSSIAN || md.kind > GSX_STRETCH_MATERN) return NAN;  // flagged as a NaN log-pdf
    if (md.kind == GSX_STRETCH_GAUSSIAN) {
        // c0 exp(-((k L / 2)^2)), c0 = (L / 2 / sqrt(pi))^d
        const double t = (r * L) / 2.0;
        s = md.c[0] * exp(-(t * t));
    } else if (md.kind == GSX_STRETCH_EXPONENTIAL) {
        // c0 / (pi (1 + (k L)^2))^((d + 1) / 2), c0 = L^d Gamma((d + 1) / 2)
        const double u = r * L;
        const double t = 3.141592653589793 * (1.0 + u * u);
        const double den = md.dim == 3 ? t * t : (md.dim == 1 ? t : pow(t, 1.5));
        s = md.c[0] / den;
    } else {
        // Matern, nu <= 20: c0 exp(((-(nu + d/2)) log(1 + x / nu) + lg(nu + d/2)) - lg(nu) - d log(sqrt(nu)))
        const double u = r * L;
        const double x = u * u;
        const double arg = ((md.c[1] * log(1.0 + x / md.c[2]) + md.c[3]) - md.c[4]) - md.c[5];
        s = md.c[0] * exp(arg);
    }
    double pdf = fac * fabs(s);
    if (!isfinite(pdf)) pdf = 0.0;
    return log(pdf);
}

// Status codes written by the kernel (first error wins).
constexpr int kStretchOk = 0;
constexpr int kStretchBadProposal = 1;  // infinite or NaN proposal
constexpr int kStretchNanLogProb = 2;   // log-probability NaN

__global__ __launch_bounds__(kWave) void k_stretch_chains(int64_t R, int nw, const uint32_t* __restrict__ states,
                                                          int64_t burn_steps, const int64_t* __restrict__ prod_steps,
                                                          int64_t max_steps, double a,
                                                          const gsx_stretch_model* __restrict__ models,
                                                          const double* __restrict__ init,
                                                          double* __restrict__ chain, double* __restrict__ final_x,
                                                          int* __restrict__ status) {
#pragma clang fp contract(off)
    __shared__ uint32_t key[kMtN];
    __shared__ uint32_t slot[kWave];
    __shared__ double x_other[kWave];
    const int64_t r = blockIdx.x;
    if (r >= R) return;
    const gsx_stretch_model md = models[r];
    const int l = lane_id();
    const bool walker = l < nw;
    double x = walker ? init[r * nw + l] : 0.0;
    double lp = walker ? ln_rad_pdf(md, x) : 0.0;
    if (__ballot(walker && isnan(lp))) {
        if (l == 0) atomicCAS(status, kStretchOk, kStretchNanLogProb);
        return;
    }
    const unsigned long long below = (l == 0) ? 0ull : (~0ull >> (64 - l));
    const unsigned long long all = nw == 64 ? ~0ull : ((1ull << nw) - 1ull);
    const int n0 = (nw + 1) / 2;
    for (int phase = 0; phase < 2; ++phase) {
        const uint32_t* st = states + (r * 2 + phase) * kStateWords;
        for (int i = l; i < kMtN; i += kWave) key[i] = st[i];
        __syncthreads();
        WaveMt mt{key, (int)st[kMtN], 0u, 0, 0};
        const int64_t steps = phase == 0 ? burn_steps : prod_steps[r];
        for (int64_t s = 0; s < steps; ++s) {
            mt.skip(2);  // the move choice (one random_sample, unused)
            unsigned long long labels = 0xaaaaaaaaaaaaaaaaull & all;  // walker w has label w % 2
            mt.shuffle_labels(labels, nw);
            for (int split = 0; split < 2; ++split) {
                const unsigned long long own = split == 0 ? (~labels & all) : labels;
                const unsigned long long oth = all & ~own;
                const int ns = split == 0 ? n0 : nw - n0;
                const int nc = nw - ns;
                const bool mine = (own >> l) & 1ull;
                const int rank = __popcll((mine ? own : oth) & below);
                if (walker && !mine) x_other[rank] = x;
                const double uz = mt.double_at(rank, mine, ns);
                mt.masked_batch((uint32_t)(nc - 1), ns, slot);
                const double ua = mt.double_at(rank, mine, ns);
                __syncthreads();
                const int p = (mine && nc > 1) ? (int)slot[rank] : 0;
                const double cp = x_other[mine ? p : 0];
                const double t = (a - 1.0) * uz + 1.0;
                const double zz = (t * t) / a;
                const double q = cp - (cp - x) * zz;
                bool bad_q = false, bad_lp = false;
                if (walker && mine) {
                    bad_q = !isfinite(q);
                    const double nl = ln_rad_pdf(md, q);
                    bad_lp = isnan(nl);
                    const double lnpdiff = (0.0 + nl) - lp;
                    if (!bad_q && !bad_lp && lnpdiff > log(ua)) {
                        x = q;
                        lp = nl;
                    }
                }
                const unsigned long long bq = __ballot(bad_q), bl = __ballot(bad_lp);
                if (bq | bl) {
                    if (l == 0) atomicCAS(status, kStretchOk, bq ? kStretchBadProposal : kStretchNanLogProb);
                    return;
                }
                __syncthreads();
            }
            if (phase == 1 && chain && walker) chain[(r * max_steps + s) * nw + l] = x;
        }
        __syncthreads();
    }
    if (final_x && walker) final_x[r * nw + l] = x;
}

// radii[o + i] = chain[r][idx[o + i]] for the picks of every chain
__global__ void k_stretch_pick(int64_t total, const int64_t* __restrict__ owner, const int64_t* __restrict__ idx,
                               int64_t max_steps, int nw, const double* __restrict__ chain, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    out[i] = chain[owner[i] * max_steps * nw + idx[i]];
}

}  // namespace gsx

extern "C" {

int gsx_stretch_run_device(int64_t R, int32_t nwalkers, const uint32_t* states, int64_t burn_steps,
                           const int64_t* prod_steps, int64_t max_steps, double a,
                           const gsx_stretch_model* models, const double* init, double* chain,
                           double* final_coords, int* status, void* stream) {
    if (R < 0 || nwalkers < 2 || nwalkers > 64 || burn_steps < 0 || max_steps < 0) return GSX_E_INVALID;
    if (R == 0) return GSX_OK;
    if (!models || !states || !prod_steps || !init || !status) return GSX_E_INVALID;
    hipStream_t s = gsx::as_stream(stream);
    hipLaunchKernelGGL(gsx::k_stretch_chains, dim3((unsigned)R), dim3(gsx::kWave), 0, s, R, (int)nwalkers, states,
                       burn_steps, prod_steps, max_steps, a, models, init, chain, final_coords, status);
    return gsx::launch_status();
}

int gsx_stretch_pick_device(int64_t total, const int64_t* owner, const int64_t* idx, int64_t max_steps,
                            int32_t nwalkers, const double* chain, double* out, void* stream) {
    if (total < 0 || max_steps < 0 || nwalkers < 1) return GSX_E_INVALID;
    if (total == 0) return GSX_OK;
    if (!owner || !idx || !chain || !out) return GSX_E_INVALID;
    hipLaunchKernelGGL(gsx::k_stretch_pick, dim3((unsigned)gsx::ceil_div(total, (int64_t)256)), dim3(256), 0,
                       gsx::as_stream(stream), total, owner, idx, max_steps, (int)nwalkers, chain, out);
    return gsx::launch_status();
}

}  // extern "C"
