// Ensemble MCMC of spectral radii on the GPU: one block of two wavefronts per
// chain.
//
// GSTools samples the radii of its spectral modes with emcee's stretch move
// (src/gstools/random/rng.py:38-104, called from generator.py:380-385).  The
// host restatement (gsx_stretch.cpp) draws every random number of a step
// from the chain's legacy numpy MT19937 stream and advances the chains in
// lockstep with one log-probability evaluation per half-move.  This file
// runs the whole chain on the device instead: block r = chain r, lane w of
// its walker wave = walker w (nwalkers <= 64), for every step of burn-in and
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
// the second: exactly the sequential recurrence.  Draws read the key at the
// stream position, each lane tempering one word; the rejection loops of the
// masked draws test 64 candidate words at once and take the first acceptable
// ones (ballot and rank), so the stream advances exactly as the sequential
// code's does.  No draw depends on the walkers' positions, so the block's
// stream wave draws step s + 1 while its walker wave applies step s.
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

// A chain runs on two wavefronts: wave 0 holds the walkers, wave 1 the
// MT19937 stream.
constexpr int kChainThreads = 2 * kWave;
// rejection hypotheses tested per ballot round of the label shuffle
constexpr int kHyp = 4;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Orders the LDS accesses of the lanes of one wavefront (a wave's LDS
// operations execute in issue order; the fences keep the compiler from
// moving accesses across this point).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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
        wave_sync();
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = lo + l + k * kWave;
            if (i < hi) key[i] = nv[k];
        }
        wave_sync();
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

    // Fisher-Yates shuffle of the labels w % 2 of nw walkers: for i = nw-1
    // down to 1, swap positions i and J[i] = masked draw in [0, i] (words
    // until (word & smear(i)) <= i).
    //
    // The draws: a window of 64 words is resolved with one ballot per
    // rejection.  With r rejections before it, window word l is tried by
    // the call i0 - (l - r), so the first lane from m on that fails under r
    // is the next rejection, or the end of the shuffle where that call index
    // drops below 1; the ballots for r .. r + kHyp - 1 are taken together.
    // Each accepted word then writes J of its call.
    //
    // The swaps, in parallel: position i is final after step i and holds
    // what position J[i] held then, which is position J[i]'s starting
    // element unless a step k > i drew J[k] = J[i] (then, for the smallest
    // such k, what position k held just before step k).  What position k
    // holds just before step k follows the same rule from k, so it is the
    // end of the chain k -> succ(k) -> ... with succ(k) the smallest k' > k
    // drawing J[k'] = k; six rounds of pointer doubling find the chain ends
    // (chains are shorter than 64), and label i is that element's label.
    __device__ unsigned long long shuffle_labels(int nw, uint32_t* jv, unsigned long long* occ) {
        const int l = lane_id();
        const unsigned long long below = (l == 0) ? 0ull : (~0ull >> (64 - l));
        int i = nw - 1;
        while (i >= 1) {
            load_window();
            const int i0 = i;
            unsigned long long rej = 0;
            int r = 0, m = 0, used = wlen;
            bool open = true;
            while (open) {
                // acceptance of every word after r + h rejections, h < kHyp
                unsigned long long acc[kHyp];
#pragma unroll
                for (int h = 0; h < kHyp; ++h) {
                    const int c = i0 - l + r + h;  // the call that tries word l
                    const uint32_t mask = c >= 1 ? (0xffffffffu >> __clz(c)) : 0u;
                    acc[h] = __ballot(l < wlen && c >= 1 && (wv & mask) <= (uint32_t)c);
                }
#pragma unroll
                for (int h = 0; h < kHyp && open; ++h) {
                    const unsigned long long ev = ~acc[h] & (~0ull << m);
                    const int e = ev ? __ffsll((long long)ev) - 1 : kWave;
                    if (e >= wlen) {  // the rest of the window accepted
                        open = false;
                    } else if (i0 - e + r < 1) {  // every call done before word e
                        used = e;
                        open = false;
                    } else {
                        rej |= 1ull << e;
                        ++r;
                        m = e + 1;
                        open = m < wlen;
                    }
                }
            }
            if (l < used && !((rej >> l) & 1ull)) {
                const int call = i0 - l + __popcll(rej & below);
                jv[call] = wv & (0xffffffffu >> __clz(call));
            }
            i = i0 - (used - __popcll(rej));
            pos = wbase + used;
        }
        occ[l] = 0ull;
        wave_sync();
        // J[0] = 0 closes position 0's rule; J[l] <= l < 64 (the mask only bounds the index)
        const int j = (l >= 1 && l < nw) ? (int)(jv[l] & (kWave - 1)) : 0;
        if (l < nw) __hip_atomic_fetch_or(&occ[j], 1ull << l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        wave_sync();
        const unsigned long long above = l == kWave - 1 ? 0ull : (~0ull << (l + 1));
        const unsigned long long same = occ[j] & above;  // later steps drawing position J[l]
        const unsigned long long onto = occ[l] & above;  // later steps drawing position l
        int end = onto ? __ffsll((long long)onto) - 1 : l;
#pragma unroll
        for (int round = 0; round < 6; ++round) end = __shfl(end, end);
        const int via = __shfl(end, same ? __ffsll((long long)same) - 1 : l);
        const int src = same ? via : j;
        wave_sync();  // the next call rewrites jv and occ
        return __ballot(l < nw && (src & 1));
    }

    // `count` masked draws in [0, m] (count <= 64): the draws are left in
    // slot[0 .. count) of LDS (read by the other wave after a barrier); m == 0
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

// The random variates of one step.  They depend on the stream alone, never
// on the walkers' positions: the shuffled labels and, per half-move, the
// stretch variate, the partner (rank among the other half) and the
// acceptance variate of the moving walker of each rank.
struct StepDraws {
    double uz[2][kWave];
    double ua[2][kWave];
    uint32_t partner[2][kWave];
    unsigned long long labels;
};

// Stream wave: every draw of one step in the sequential order (the move
// choice, the label shuffle, then per half-move ns stretch doubles, ns
// partner draws and ns acceptance doubles).
__device__ void draw_step(WaveMt& mt, int nw, int n0, StepDraws& d, uint32_t* jv, unsigned long long* occ) {
#pragma clang fp contract(off)
    const int l = lane_id();
    mt.skip(2);  // the move choice (one random_sample, unused)
    const unsigned long long labels = mt.shuffle_labels(nw, jv, occ);
    if (l == 0) d.labels = labels;
    for (int split = 0; split < 2; ++split) {
        const int ns = split == 0 ? n0 : nw - n0;  // walkers with label `split`
        const int nc = nw - ns;
        const bool need = l < ns;
        const double uz = mt.double_at(l, need, ns);
        mt.masked_batch((uint32_t)(nc - 1), ns, d.partner[split]);
        const double ua = mt.double_at(l, need, ns);
        if (need) {
            d.uz[split][l] = uz;
            d.ua[split][l] = ua;
        }
    }
}

// Walker wave: the two half-moves of one step from its draws.  Returns false
// (status set) on an infinite or NaN proposal or a NaN log-probability.
__device__ bool apply_step(const StepDraws& d, const gsx_stretch_model& md, int nw, int n0,
                           unsigned long long all, double a, double& x, double& lp, double* x_other,
                           int* status) {
#pragma clang fp contract(off)
    const int l = lane_id();
    const bool walker = l < nw;
    const unsigned long long below = (l == 0) ? 0ull : (~0ull >> (64 - l));
    const unsigned long long labels = d.labels;
    for (int split = 0; split < 2; ++split) {
        const unsigned long long own = split == 0 ? (~labels & all) : labels;
        const unsigned long long oth = all & ~own;
        const int ns = split == 0 ? n0 : nw - n0;
        const int nc = nw - ns;
        const bool mine = (own >> l) & 1ull;
        const int rank = __popcll((mine ? own : oth) & below);
        if (walker && !mine) x_other[rank] = x;
        wave_sync();
        const int p = (mine && nc > 1) ? (int)d.partner[split][rank] : 0;
        const double cp = x_other[mine ? p : 0];
        const double uz = mine ? d.uz[split][rank] : 0.0;
        const double t = (a - 1.0) * uz + 1.0;
        const double zz = (t * t) / a;
        const double q = cp - (cp - x) * zz;
        bool bad_q = false, bad_lp = false;
        if (walker && mine) {
            bad_q = !isfinite(q);
            const double nl = ln_rad_pdf(md, q);
            bad_lp = isnan(nl);
            const double lnpdiff = (0.0 + nl) - lp;
            if (!bad_q && !bad_lp && lnpdiff > log(d.ua[split][rank])) {
                x = q;
                lp = nl;
            }
        }
        const unsigned long long bq = __ballot(bad_q), bl = __ballot(bad_lp);
        if (bq | bl) {
            if (l == 0) atomicCAS(status, kStretchOk, bq ? kStretchBadProposal : kStretchNanLogProb);
            return false;
        }
        wave_sync();  // the next half-move rewrites x_other
    }
    return true;
}

// One block of two waves per chain.  The stream wave draws step s while the
// walker wave applies step s - 1, double-buffered in LDS with one barrier
// per step; both waves run the same loop, so every barrier is reached by
// both, and a stop flag raised by the walker wave ends both after the
// barrier.
__global__ __launch_bounds__(kChainThreads) void k_stretch_chains(int64_t R, int nw,
                                                                  const uint32_t* __restrict__ states,
                                                                  int64_t burn_steps,
                                                                  const int64_t* __restrict__ prod_steps,
                                                                  int64_t max_steps, double a,
                                                                  const gsx_stretch_model* __restrict__ models,
                                                                  const double* __restrict__ init,
                                                                  double* __restrict__ chain,
                                                                  double* __restrict__ final_x,
                                                                  int* __restrict__ status) {
#pragma clang fp contract(off)
    __shared__ uint32_t key[kMtN];
    __shared__ double x_other[kWave];
    __shared__ StepDraws draws[2];
    __shared__ uint32_t shuffle_j[kWave];
    __shared__ unsigned long long shuffle_occ[kWave];
    __shared__ int stop;
    const int64_t r = blockIdx.x;
    if (r >= R) return;
    const gsx_stretch_model md = models[r];
    // the wave's role, uniform across its lanes (scalar branches and loops)
    const bool stream_wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x) >= kWave;
    const int l = lane_id();
    const bool walker = !stream_wave && l < nw;
    double x = walker ? init[r * nw + l] : 0.0;
    double lp = walker ? ln_rad_pdf(md, x) : 0.0;
    if (threadIdx.x == 0) stop = 0;
    if (!stream_wave && __ballot(walker && isnan(lp))) {
        if (l == 0) {
            atomicCAS(status, kStretchOk, kStretchNanLogProb);
            stop = 1;
        }
    }
    __syncthreads();
    if (stop) return;
    const unsigned long long all = nw == 64 ? ~0ull : ((1ull << nw) - 1ull);
    const int n0 = (nw + 1) / 2;
    for (int phase = 0; phase < 2; ++phase) {
        const int64_t steps = phase == 0 ? burn_steps : prod_steps[r];
        WaveMt mt{key, 0, 0u, 0, 0};
        if (stream_wave) {
            const uint32_t* st = states + (r * 2 + phase) * kStateWords;
            for (int i = l; i < kMtN; i += kWave) key[i] = st[i];
            mt.pos = (int)st[kMtN];
            wave_sync();
        }
        for (int64_t s = 0; s <= steps; ++s) {
            if (stream_wave) {
                if (s < steps) draw_step(mt, nw, n0, draws[s & 1], shuffle_j, shuffle_occ);
            } else if (s > 0) {
                if (!apply_step(draws[(s - 1) & 1], md, nw, n0, all, a, x, lp, x_other, status)) {
                    if (l == 0) stop = 1;
                } else if (phase == 1 && chain && walker) {
                    chain[(r * max_steps + s - 1) * nw + l] = x;
                }
            }
            __syncthreads();
            if (stop) return;
        }
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
    hipLaunchKernelGGL(gsx::k_stretch_chains, dim3((unsigned)R), dim3(gsx::kChainThreads), 0, s, R, (int)nwalkers, states,
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
