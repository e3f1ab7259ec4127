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
// 64-word chunks in ascending order: word i reads words i + 1 (the next
// chunk's first word is still old) and (i + 397) mod 624 (old for i < 227,
// already new for i >= 227, at least 227 > 64 positions back), exactly the
// sequential recurrence.  Draws read the key at the stream position, each
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

// One chain's MT19937 stream, shared by the 64 lanes of its wave.
struct WaveMt {
    uint32_t* key;  // LDS, kMtN words
    int pos;        // stream position (uniform across the wave)

    __device__ void twist() {
        const int l = lane_id();
        for (int i0 = 0; i0 < kMtN; i0 += kWave) {
            const int i = i0 + l;
            uint32_t nv = 0;
            if (i < kMtN) {
                const uint32_t y = (key[i] & kUpper) | (key[i + 1 < kMtN ? i + 1 : 0] & kLower);
                const int far = i + kMtM < kMtN ? i + kMtM : i + kMtM - kMtN;
                nv = key[far] ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
            }
            __syncthreads();
            if (i < kMtN) key[i] = nv;
            __syncthreads();
        }
        pos = 0;
    }

    // lanes 0..n-1 receive the next n words (n <= 64)
    __device__ uint32_t take(int n) {
        const int l = lane_id();
        const int avail = kMtN - pos;
        uint32_t v = 0;
        if (l < n && l < avail) v = temper(key[pos + l]);
        if (n > avail) {
            twist();
            if (l >= avail && l < n) v = temper(key[l - avail]);
            pos = n - avail;
        } else {
            pos += n;
        }
        return v;
    }

    // random_sample(): lane i < count receives the i-th of `count` doubles (count <= 32)
    __device__ double doubles(int count) {
        const uint32_t w = take(2 * count);
        const int l = lane_id();
        const int src = 2 * (l < count ? l : 0);
        const uint32_t a = (uint32_t)__shfl((int)w, src) >> 5;
        const uint32_t b = (uint32_t)__shfl((int)w, src + 1) >> 6;
        return ((double)a * 67108864.0 + (double)b) / 9007199254740992.0;
    }

    // one uniform integer in [0, m] by masked rejection; the same value in every lane
    __device__ uint32_t masked_one(uint32_t m) {
        if (m == 0) return 0;
        const uint32_t mask = smear(m);
        const int l = lane_id();
        for (;;) {
            int avail = kMtN - pos;
            if (avail == 0) {
                twist();
                avail = kMtN;
            }
            const int n = avail < kWave ? avail : kWave;
            const uint32_t v = l < n ? (temper(key[pos + l]) & mask) : 0xffffffffu;
            const unsigned long long ok = __ballot(l < n && v <= m);
            if (ok) {
                const int f = __ffsll((long long)ok) - 1;
                pos += f + 1;
                return (uint32_t)__shfl((int)v, f);
            }
            pos += n;
        }
    }

    // `count` masked draws in [0, m] (count <= 64): lane i receives the i-th;
    // slot is a 64-entry LDS scratch row
    __device__ uint32_t masked_batch(uint32_t m, int count, uint32_t* slot) {
        if (m == 0) return 0;
        const uint32_t mask = smear(m);
        const int l = lane_id();
        const unsigned long long below = (l == 0) ? 0ull : (~0ull >> (64 - l));
        int got = 0;
        while (got < count) {
            int avail = kMtN - pos;
            if (avail == 0) {
                twist();
                avail = kMtN;
            }
            const int n = avail < kWave ? avail : kWave;
            const uint32_t v = l < n ? (temper(key[pos + l]) & mask) : 0xffffffffu;
            const bool ok = l < n && v <= m;
            const unsigned long long bal = __ballot(ok);
            const int rank = __popcll(bal & below);
            const int need = count - got;
            if (ok && rank < need) slot[got + rank] = v;
            const int total = __popcll(bal);
            if (total >= need) {
                const unsigned long long last = __ballot(ok && rank == need - 1);
                pos += __ffsll((long long)last);  // index of that lane + 1
                got = count;
            } else {
                pos += n;
                got += total;
            }
        }
        __syncthreads();
        const uint32_t out = l < count ? slot[l] : 0u;
        __syncthreads();
        return out;
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
                                                          int64_t max_steps, double a, gsx_stretch_model md,
                                                          const double* __restrict__ init,
                                                          double* __restrict__ chain, double* __restrict__ final_x,
                                                          int* __restrict__ status) {
#pragma clang fp contract(off)
    __shared__ uint32_t key[kMtN];
    __shared__ uint32_t slot[kWave];
    __shared__ int other_of[kWave];
    const int64_t r = blockIdx.x;
    if (r >= R) return;
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
        WaveMt mt{key, (int)st[kMtN]};
        const int64_t steps = phase == 0 ? burn_steps : prod_steps[r];
        for (int64_t s = 0; s < steps; ++s) {
            (void)mt.take(2);  // move choice
            unsigned long long labels = 0xaaaaaaaaaaaaaaaaull & all;  // walker w has label w % 2
            for (int i = nw - 1; i >= 1; --i) {
                const uint32_t j = mt.masked_one((uint32_t)i);
                const unsigned long long bi = (labels >> i) & 1ull, bj = (labels >> j) & 1ull;
                if (bi != bj) labels ^= (1ull << i) | (1ull << j);
            }
            for (int split = 0; split < 2; ++split) {
                const unsigned long long own = split == 0 ? (~labels & all) : labels;
                const unsigned long long oth = all & ~own;
                const int ns = split == 0 ? n0 : nw - n0;
                const int nc = nw - ns;
                const bool mine = (own >> l) & 1ull;
                const int rank = __popcll((mine ? own : oth) & below);
                if (walker && !mine) other_of[rank] = l;
                const double u_z = mt.doubles(ns);
                const uint32_t part = mt.masked_batch((uint32_t)(nc - 1), ns, slot);
                const double u_acc = mt.doubles(ns);
                __syncthreads();
                // walker of rank i in `own` takes draw i of each batch
                const int src = mine ? rank : 0;
                const double uz = __shfl(u_z, src);
                const int p = (int)__shfl((int)part, src);
                const double ua = __shfl(u_acc, src);
                const int pw = other_of[mine ? p : 0];
                const double cp = __shfl(x, pw);
                const double t = (a - 1.0) * uz + 1.0;
                const double zz = (t * t) / a;
                const double q = cp - (cp - x) * zz;
                bool bad_q = false, bad_lp = false;
                double nl = 0.0;
                if (walker && mine) {
                    bad_q = !isfinite(q);
                    nl = ln_rad_pdf(md, q);
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
                           const int64_t* prod_steps, int64_t max_steps, double a, const gsx_stretch_model* model,
                           const double* init, double* chain, double* final_coords, int* status, void* stream) {
    if (R < 0 || nwalkers < 2 || nwalkers > 64 || burn_steps < 0 || max_steps < 0 || !model) return GSX_E_INVALID;
    if (model->kind < GSX_STRETCH_GAUSSIAN || model->kind > GSX_STRETCH_MATERN || model->dim < 1 ||
        model->dim > 3)
        return GSX_E_INVALID;
    if (R == 0) return GSX_OK;
    if (!states || !prod_steps || !init || !status) return GSX_E_INVALID;
    hipStream_t s = gsx::as_stream(stream);
    hipLaunchKernelGGL(gsx::k_stretch_chains, dim3((unsigned)R), dim3(gsx::kWave), 0, s, R, (int)nwalkers, states,
                       burn_steps, prod_steps, max_steps, a, *model, init, chain, final_coords, status);
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
