// Host-side random draws of the affine-invariant ensemble sampler, for many
// chains at once.
//
// GSTools samples the radii of its spectral modes with emcee's stretch move
// (src/gstools/random/rng.py:38-104, called from generator.py:380-385); every
// random number comes from a legacy numpy RandomState (Mersenne Twister
// MT19937).  None of the draws depends on the state of the walkers, so all
// draws of a step can be made before the step's two half-moves run.  This
// file makes them for R independent chains in one call; the host code then
// advances the R chains in lockstep with one probability evaluation per
// half-move (gstools_amd/sampling.py).
//
// The same file holds the two halves of a half-move (propose, accept) for
// one-dimensional walkers, so that only the model's log-probability is
// evaluated outside native code.  Their arithmetic is emcee's, rounded the
// same way (every product and sum rounded, no fused multiply-add):
//   zz = ((a - 1) u + 1)^2 / a,   q = c_p - (c_p - x) zz,
//   accept iff (0 + lp(q)) - lp(x) > log(u_acc)
// where the log-stretch term (ndim - 1) log(zz) is 0 for ndim = 1 and
// log(u_acc) is supplied by the caller.
//
// Per step and chain, in the order the sampler consumes the stream:
//   1. one random_sample() (the choice of the single move; discarded);
//   2. shuffle of the labels arange(nwalkers) % 2: Fisher-Yates from the last
//      position down, each swap partner j in [0, i] drawn by the masked
//      rejection "interval" draw;
//   3. for label 0, then label 1 (ns walkers carry the label, nc do not):
//      ns random_sample() stretch variates, ns masked draws in [0, nc - 1]
//      for the partner walkers, ns random_sample() for the acceptance test.
//
// The generator is the published MT19937 (Matsumoto & Nishimura, ACM TOMACS
// 8(1), 1998): 624-word state, twist with matrix 0x9908b0df, tempering
// (11, 7/0x9d2c5680, 15/0xefc60000, 18).  random_sample() joins two outputs
// into a 53-bit fraction, (a >> 5) * 2^26 + (b >> 6) over 2^53; the masked
// draw in [0, m] keeps the low bits of one output under the smallest
// all-ones mask >= m and rejects values above m.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/gsx.h"

namespace {

constexpr int kMtN = 624;
constexpr int kMtM = 397;
constexpr uint32_t kMatrixA = 0x9908b0dfu;
constexpr uint32_t kUpper = 0x80000000u;
constexpr uint32_t kLower = 0x7fffffffu;

struct Mt {
    uint32_t* key;  // kMtN words followed by the position word

    uint32_t& pos() { return key[kMtN]; }

    static uint32_t mix(uint32_t hi, uint32_t lo, uint32_t far) {
        const uint32_t y = (hi & kUpper) | (lo & kLower);
        return far ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
    }

    void twist() {
        int i = 0;
        for (; i < kMtN - kMtM; ++i) key[i] = mix(key[i], key[i + 1], key[i + kMtM]);
        for (; i < kMtN - 1; ++i) key[i] = mix(key[i], key[i + 1], key[i + kMtM - kMtN]);
        key[kMtN - 1] = mix(key[kMtN - 1], key[0], key[kMtM - 1]);
        pos() = 0;
    }

    uint32_t next32() {
        if (pos() >= (uint32_t)kMtN) twist();
        uint32_t y = key[pos()++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }

    double next_double() {
        const uint32_t a = next32() >> 5, b = next32() >> 6;
        return (a * 67108864.0 + b) / 9007199254740992.0;
    }

    // uniform integer in [0, m] by masked rejection (m < 2^32)
    uint32_t masked(uint32_t m) {
        if (m == 0) return 0;
        uint32_t mask = m;
        mask |= mask >> 1;
        mask |= mask >> 2;
        mask |= mask >> 4;
        mask |= mask >> 8;
        mask |= mask >> 16;
        uint32_t v;
        while ((v = (next32() & mask)) > m) {
        }
        return v;
    }
};

// A legacy numpy RandomState: the MT19937 stream plus the cached second
// normal of the polar method (numpy/random legacy distributions).
struct Legacy {
    uint32_t words[kMtN + 1];
    bool has_gauss = false;
    double gauss = 0.0;

    // RandomState(seed) for an integer seed in [0, 2^32): Knuth's
    // initialisation of the key (multiplier 1812433253), position at the end
    void seed(uint32_t s) {
        words[0] = s;
        for (int i = 1; i < kMtN; ++i) words[i] = 1812433253u * (words[i - 1] ^ (words[i - 1] >> 30)) + (uint32_t)i;
        words[kMtN] = kMtN;
        has_gauss = false;
        gauss = 0.0;
    }
    Mt mt() { return Mt{words}; }
    double random_sample() { return mt().next_double(); }
    // randint(low, high) for high - low <= 2^32: low + masked draw in [0, high - low - 1]
    int64_t randint(int64_t low, int64_t high) { return low + (int64_t)mt().masked((uint32_t)(high - low - 1)); }
    // standard normal by the polar method, every product and sum rounded
    double gauss_draw() {
        if (has_gauss) {
            has_gauss = false;
            const double t = gauss;
            gauss = 0.0;
            return t;
        }
        double x1, x2, r2;
        do {
            x1 = 2.0 * random_sample() - 1.0;
            x2 = 2.0 * random_sample() - 1.0;
            r2 = x1 * x1 + x2 * x2;
        } while (r2 >= 1.0 || r2 == 0.0);
        const double f = std::sqrt(-2.0 * std::log(r2) / r2);
        gauss = f * x1;
        has_gauss = true;
        return f * x2;
    }
    double normal() { return 0.0 + 1.0 * gauss_draw(); }            // normal(0, 1)
    double uniform(double low, double range) { return low + range * random_sample(); }
};

void chain_draws(Mt mt, int64_t steps, int32_t nw, int8_t* labels, double* uz, int64_t* partner,
                 double* uacc) {
    const int32_t n0 = (nw + 1) / 2;  // walkers labelled 0
    for (int64_t s = 0; s < steps; ++s) {
        int8_t* lab = labels + s * nw;
        double* z = uz + s * nw;
        int64_t* pr = partner + s * nw;
        double* ua = uacc + s * nw;
        (void)mt.next_double();  // move choice
        for (int32_t w = 0; w < nw; ++w) lab[w] = (int8_t)(w % 2);
        for (int32_t i = nw - 1; i >= 1; --i) {
            const uint32_t j = mt.masked((uint32_t)i);
            const int8_t t = lab[i];
            lab[i] = lab[j];
            lab[j] = t;
        }
        for (int split = 0; split < 2; ++split) {
            const int32_t ns = split == 0 ? n0 : nw - n0;
            const int32_t nc = nw - ns;
            const int32_t off = split == 0 ? 0 : n0;
            for (int32_t k = 0; k < ns; ++k) z[off + k] = mt.next_double();
            for (int32_t k = 0; k < ns; ++k) pr[off + k] = (int64_t)mt.masked((uint32_t)(nc - 1));
            for (int32_t k = 0; k < ns; ++k) ua[off + k] = mt.next_double();
        }
    }
}

// Walkers of chain row `lab` carrying label `split` (in walker order) and the others.
inline void split_walkers(const int8_t* lab, int32_t nw, int32_t split, int32_t* own, int32_t* other) {
    int32_t a = 0, b = 0;
    for (int32_t w = 0; w < nw; ++w) {
        if (lab[w] == split)
            own[a++] = w;
        else
            other[b++] = w;
    }
}

}  // namespace

extern "C" {

int gsx_stretch_draws(int64_t R, uint32_t* states, const int64_t* steps, int64_t max_steps, int32_t nwalkers,
                      int8_t* labels, double* uz, int64_t* partner, double* uacc, int32_t nthreads) {
    if (R < 0 || max_steps < 0 || nwalkers < 2) return GSX_E_INVALID;
    if (R == 0 || max_steps == 0) return GSX_OK;
    if (!states || !steps || !labels || !uz || !partner || !uacc) return GSX_E_INVALID;
    for (int64_t r = 0; r < R; ++r)
        if (steps[r] < 0 || steps[r] > max_steps || states[r * (kMtN + 1) + kMtN] > (uint32_t)kMtN)
            return GSX_E_INVALID;
    const int64_t per_chain = max_steps * nwalkers;
    auto run = [&](int64_t r) {
        Mt mt{states + r * (kMtN + 1)};
        chain_draws(mt, steps[r], nwalkers, labels + r * per_chain, uz + r * per_chain, partner + r * per_chain,
                    uacc + r * per_chain);
    };
    int64_t workers = nthreads > 0 ? nthreads : 1;
    if (workers > R) workers = R;
    if (workers <= 1) {
        for (int64_t r = 0; r < R; ++r) run(r);
        return GSX_OK;
    }
    std::vector<std::thread> pool;
    pool.reserve((size_t)workers);
    for (int64_t t = 0; t < workers; ++t)
        pool.emplace_back([&, t] {
            for (int64_t r = t; r < R; r += workers) run(r);
        });
    for (auto& th : pool) th.join();
    return GSX_OK;
}

int gsx_mode_draws(int64_t R, const int64_t* seeds, const int64_t* n_modes, int32_t dim, int32_t kind,
                   int32_t nwalkers, double sample_around, const int64_t* n_flat, double* z1, double* z2,
                   double* ang1, double* ang2, double* u_inv, double* init, uint32_t* states, int64_t* picks,
                   int32_t nthreads) {
    if (R < 0 || dim < 1 || dim > 3 || kind < 0 || kind > 2) return GSX_E_INVALID;
    if (R == 0) return GSX_OK;
    if (!seeds || !n_modes || !z1 || !z2) return GSX_E_INVALID;
    if (kind > 0 && !ang1) return GSX_E_INVALID;
    if (kind > 0 && dim == 3 && !ang2) return GSX_E_INVALID;
    if (kind == 1 && !u_inv) return GSX_E_INVALID;
    if (kind == 2 && (nwalkers < 1 || !n_flat || !init || !states || !picks)) return GSX_E_INVALID;
    std::vector<int64_t> off((size_t)R + 1, 0);
    for (int64_t r = 0; r < R; ++r) {
        if (seeds[r] < 0 || seeds[r] > 0xffffffffLL || n_modes[r] < 0) return GSX_E_INVALID;
        if (kind == 2 && (n_flat[r] < 1 || n_flat[r] > 0xffffffffLL)) return GSX_E_INVALID;
        off[(size_t)r + 1] = off[(size_t)r] + n_modes[r];
    }
    const double two_pi = 6.283185307179586;  // 2 * pi - 0.0, the range of uniform(0, 2 pi)
    auto run = [&](int64_t r) {
        Legacy master, stream;
        master.seed((uint32_t)seeds[r]);
        auto next_stream = [&]() { stream.seed((uint32_t)master.randint(1, 65536)); };
        const int64_t n = n_modes[r], o = off[(size_t)r];
        next_stream();
        for (int64_t i = 0; i < n; ++i) z1[o + i] = stream.normal();
        next_stream();
        for (int64_t i = 0; i < n; ++i) z2[o + i] = stream.normal();
        if (kind == 0) return;
        next_stream();
        if (dim == 1) {  // choice([-1, 1]): randint(0, 2) picks the entry
            for (int64_t i = 0; i < n; ++i) ang1[o + i] = stream.randint(0, 2) == 0 ? -1.0 : 1.0;
        } else {
            for (int64_t i = 0; i < n; ++i) ang1[o + i] = stream.uniform(0.0, two_pi);
            if (dim == 3) {
                next_stream();
                for (int64_t i = 0; i < n; ++i) ang2[o + i] = stream.uniform(-1.0, 2.0);
            }
        }
        if (kind == 1) {  // inverse-cdf sampling: uniform(size=n)
            next_stream();
            for (int64_t i = 0; i < n; ++i) u_inv[o + i] = stream.uniform(0.0, 1.0);
            return;
        }
        next_stream();  // initial walkers: rand(nwalkers) * sample_around
        for (int32_t w = 0; w < nwalkers; ++w) init[r * nwalkers + w] = stream.random_sample() * sample_around;
        for (int k = 0; k < 2; ++k) {  // burn-in and production streams, fresh
            next_stream();
            uint32_t* dst = states + (r * 2 + k) * (kMtN + 1);
            for (int i = 0; i <= kMtN; ++i) dst[i] = stream.words[i];
        }
        next_stream();  // the final choice(n_flat, n)
        for (int64_t i = 0; i < n; ++i) picks[o + i] = stream.randint(0, n_flat[r]);
    };
    int64_t workers = nthreads > 0 ? nthreads : 1;
    if (workers > R) workers = R;
    if (workers <= 1) {
        for (int64_t r = 0; r < R; ++r) run(r);
        return GSX_OK;
    }
    std::vector<std::thread> pool;
    pool.reserve((size_t)workers);
    for (int64_t t = 0; t < workers; ++t)
        pool.emplace_back([&, t] {
            for (int64_t r = t; r < R; r += workers) run(r);
        });
    for (auto& th : pool) th.join();
    return GSX_OK;
}

int gsx_stretch_propose(int64_t R, int32_t nwalkers, const int64_t* steps, int64_t step, int64_t block_row,
                        int64_t block, int32_t split, double a, const double* coords, const int8_t* labels,
                        const double* uz, const int64_t* partner, double* q, int64_t* n_active) {
    if (R < 0 || nwalkers < 2 || nwalkers > 4096 || (split != 0 && split != 1) || block_row < 0 ||
        block_row >= block || !n_active)
        return GSX_E_INVALID;
    if (R > 0 && (!steps || !coords || !labels || !uz || !partner || !q)) return GSX_E_INVALID;
    const int32_t n0 = (nwalkers + 1) / 2;
    const int32_t ns = split == 0 ? n0 : nwalkers - n0;
    const int32_t nc = nwalkers - ns;
    const int32_t off = split == 0 ? 0 : n0;
    int32_t own[4096], other[4096];
    int64_t k = 0;
    for (int64_t r = 0; r < R; ++r) {
        if (steps[r] <= step) continue;
        const int64_t d = (r * block + block_row) * nwalkers;
        split_walkers(labels + d, nwalkers, split, own, other);
        const double* x = coords + r * nwalkers;
        for (int32_t i = 0; i < ns; ++i) {
            const int64_t p = partner[d + off + i];
            if (p < 0 || p >= nc) return GSX_E_INVALID;
            const double t = (a - 1.0) * uz[d + off + i] + 1.0;
            const double zz = (t * t) / a;
            const double cp = x[other[p]];
            q[k * ns + i] = cp - (cp - x[own[i]]) * zz;
        }
        ++k;
    }
    *n_active = k;
    return GSX_OK;
}

int gsx_stretch_accept(int64_t R, int32_t nwalkers, const int64_t* steps, int64_t step, int64_t block_row,
                       int64_t block, int32_t split, const int8_t* labels, const double* log_uacc,
                       const double* q, const double* new_lp, double* coords, double* log_prob, double* chain,
                       int64_t chain_steps) {
    if (R < 0 || nwalkers < 2 || nwalkers > 4096 || (split != 0 && split != 1) || block_row < 0 ||
        block_row >= block)
        return GSX_E_INVALID;
    if (R > 0 && (!steps || !labels || !log_uacc || !q || !new_lp || !coords || !log_prob)) return GSX_E_INVALID;
    if (chain && step >= chain_steps) return GSX_E_INVALID;
    const int32_t n0 = (nwalkers + 1) / 2;
    const int32_t ns = split == 0 ? n0 : nwalkers - n0;
    const int32_t off = split == 0 ? 0 : n0;
    int32_t own[4096], other[4096];
    int64_t k = 0;
    for (int64_t r = 0; r < R; ++r) {
        if (steps[r] <= step) continue;
        const int64_t d = (r * block + block_row) * nwalkers;
        split_walkers(labels + d, nwalkers, split, own, other);
        double* x = coords + r * nwalkers;
        double* lp = log_prob + r * nwalkers;
        for (int32_t i = 0; i < ns; ++i) {
            const double nl = new_lp[k * ns + i];
            const double lnpdiff = (0.0 + nl) - lp[own[i]];
            if (lnpdiff > log_uacc[d + off + i]) {
                x[own[i]] = q[k * ns + i];
                lp[own[i]] = nl;
            }
        }
        if (chain)
            for (int32_t w = 0; w < nwalkers; ++w) chain[(r * chain_steps + step) * nwalkers + w] = x[w];
        ++k;
    }
    return GSX_OK;
}

}  // extern "C"
