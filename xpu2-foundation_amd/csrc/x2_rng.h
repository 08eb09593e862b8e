// x2_rng.h — the sample-identity RNG on the device (gfx950): Philox4x32-10 keyed by (stream, branch), counters
// (pixel, sample index, dimension group), and Owen-scrambled Sobol for the policy dimensions. Every random number
// a sample draws is a pure function of its identity, so the device reproduces the reference's stream word for word
// (xpu2/interactive/sample_stream.h:25-162; paths::Stream and childBranch, async_integrator.h:151-176,281-286).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace x2 {

// Sobol direction numbers, 17 dimensions x 32 bits (sample_stream.h:82-100), generated on the host from the
// Joe-Kuo polynomials (x2_lane.hip) and uploaded once per device.
extern __constant__ uint32_t kSobolDirections[17 * 32];

struct Words { uint32_t v[4]; };

// a ^ b ^ c in one instruction (gfx950's three-input bitwise op, truth table 0x96)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Philox4x32-10 (sample_stream.h:31-38): 10 rounds of two 32x32->64 multiplies and a Weyl key schedule. Each
// product is one 64-bit multiply (v_mad_u64_u32), its halves the round's high and low words.
__device__ __forceinline__ Words philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t ka, uint32_t kb) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t pa = uint64_t(c0) * 0xd2511f53u, pb = uint64_t(c2) * 0xcd9e8d57u;
        const uint32_t alo = uint32_t(pa), ahi = uint32_t(pa >> 32), blo = uint32_t(pb), bhi = uint32_t(pb >> 32);
        c0 = xor3(bhi, c1, ka);
        c1 = blo;
        c2 = xor3(ahi, c3, kb);
        c3 = alo;
        ka += 0x9e3779b9u;
        kb += 0xbb67ae85u;
    }
    return Words{{c0, c1, c2, c3}};
}

// A sample's identity (sample_stream.h:42): pixel, 64-bit sample index, frame stream and 64-bit branch address.
struct Identity {
    uint32_t pixel;
    uint64_t index, stream, branch;
};

// block (sample_stream.h:43-51): key = stream ^ branch * golden ratio; counter = (pixel, index lo, index hi, group).
__device__ __forceinline__ Words block(const Identity& id, uint32_t group) {
    const uint64_t k = id.stream ^ (id.branch * 0x9e3779b97f4a7c15ull);
    return philox(id.pixel, uint32_t(id.index), uint32_t(id.index >> 32), group, uint32_t(k), uint32_t(k >> 32));
}
// Two blocks computed together, their rounds interleaved: four independent multiplies a round instead of two, so a
// round's latency is spent once for both (the same words as two block() calls).
__device__ __forceinline__ void block2(const Identity& a, uint32_t ga, const Identity& b, uint32_t gb, Words& wa,
                                       Words& wb) {
    const uint64_t ka = a.stream ^ (a.branch * 0x9e3779b97f4a7c15ull), kb = b.stream ^ (b.branch * 0x9e3779b97f4a7c15ull);
    uint32_t a0 = a.pixel, a1 = uint32_t(a.index), a2 = uint32_t(a.index >> 32), a3 = ga, aka = uint32_t(ka),
             akb = uint32_t(ka >> 32);
    uint32_t b0 = b.pixel, b1 = uint32_t(b.index), b2 = uint32_t(b.index >> 32), b3 = gb, bka = uint32_t(kb),
             bkb = uint32_t(kb >> 32);
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t pa = uint64_t(a0) * 0xd2511f53u, pb = uint64_t(a2) * 0xcd9e8d57u;
        const uint64_t qa = uint64_t(b0) * 0xd2511f53u, qb = uint64_t(b2) * 0xcd9e8d57u;
        a0 = xor3(uint32_t(pb >> 32), a1, aka);
        a1 = uint32_t(pb);
        a2 = xor3(uint32_t(pa >> 32), a3, akb);
        a3 = uint32_t(pa);
        b0 = xor3(uint32_t(qb >> 32), b1, bka);
        b1 = uint32_t(qb);
        b2 = xor3(uint32_t(qa >> 32), b3, bkb);
        b3 = uint32_t(qa);
        aka += 0x9e3779b9u;
        akb += 0xbb67ae85u;
        bka += 0x9e3779b9u;
        bkb += 0xbb67ae85u;
    }
    wa = Words{{a0, a1, a2, a3}};
    wb = Words{{b0, b1, b2, b3}};
}
__device__ __forceinline__ uint32_t pick(const Words& w, uint32_t i) {
    return i == 0 ? w.v[0] : i == 1 ? w.v[1] : i == 2 ? w.v[2] : w.v[3];
}
__device__ __forceinline__ uint32_t word(const Identity& id, uint32_t dimension) {
    return pick(block(id, dimension >> 2), dimension & 3);
}
// 24 bits times 2^-24: exact, in [0, 1) (sample_stream.h:56).
__device__ __forceinline__ float unit(uint32_t bits) { return float(bits >> 8) * 0x1p-24f; }

__device__ __forceinline__ uint32_t sobol(uint32_t index, uint32_t dimension) {  // sample_stream.h:102-107
    uint32_t gray = index ^ (index >> 1), result = 0, bit = 0;
    const uint32_t* d = kSobolDirections + dimension * 32;
    while (gray) {
        if (gray & 1u) result ^= d[bit];
        gray >>= 1;
        ++bit;
    }
    return result;
}
// Burley's hash-based Owen scramble (JCGT 2020; sample_stream.h:122-126), bit reversal by the hardware.
__device__ __forceinline__ uint32_t owenScramble(uint32_t x, uint32_t seed) {
    x = __brev(x);
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return __brev(x);
}
// The policy dimensions (sample_stream.h:127-144): camera 0..4, first-hit BSDF 16/17, first-hit light 18/19,
// subsurface 20..27 take Owen-scrambled Sobol when their policy bit is set; everything else is plain Philox.
__device__ __forceinline__ uint32_t samplingWord(const Identity& id, uint32_t dimension, uint32_t policy) {
    const bool camera = dimension < 5, diffuse = dimension == 16 || dimension == 17,
               light = dimension == 18 || dimension == 19, subsurface = dimension >= 20 && dimension <= 27;
    if ((!camera || !(policy & 1)) && (!diffuse || !(policy & 2)) && (!light || !(policy & 4)) &&
        (!subsurface || !(policy & 8)))
        return word(id, dimension);
    const uint32_t d = camera ? (dimension < 2 ? dimension : dimension < 4 ? dimension + 12 : 16) : dimension - 14;
    Identity shift = id;
    shift.index = id.index >> 32;
    if (camera && dimension >= 2) {
        const Words seeds = block(shift, dimension < 4 ? 0x50414444u : 0x50414445u);
        return owenScramble(sobol(owenScramble(uint32_t(id.index), seeds.v[0]), dimension < 4 ? dimension - 2 : 0u),
                            pick(seeds, dimension < 4 ? dimension - 1 : 1));
    }
    return owenScramble(sobol(uint32_t(id.index), d), pick(block(shift, 0x534f424fu + (d >> 2)), d & 3));
}
__device__ __forceinline__ float draw(const Identity& id, uint32_t dimension, uint32_t policy) {
    return unit(samplingWord(id, dimension, policy));
}
// drawPair in two steps, for a caller that computes the pair's block together with another: pairSource names the
// block (identity and group) the pair reads, pairFrom makes the two draws from it.
__device__ __forceinline__ void pairSource(const Identity& id, uint32_t dimension, uint32_t policy, Identity& src,
                                           uint32_t& group) {
    const bool camera = dimension < 5, diffuse = dimension == 16, light = dimension == 18;
    const bool sobolPair = (camera && (policy & 1)) || (diffuse && (policy & 2)) || (light && (policy & 4));
    src = id;
    if (!sobolPair) {
        group = dimension >> 2;
        return;
    }
    const uint32_t d = camera ? dimension : dimension - 14;
    src.index = id.index >> 32;
    group = 0x534f424fu + (d >> 2);
}
__device__ __forceinline__ void pairFrom(const Identity& id, uint32_t dimension, uint32_t policy, const Words& w,
                                         float& u0, float& u1) {
    const bool camera = dimension < 5, diffuse = dimension == 16, light = dimension == 18;
    const bool sobolPair = (camera && (policy & 1)) || (diffuse && (policy & 2)) || (light && (policy & 4));
    if (!sobolPair) {
        u0 = unit(pick(w, dimension & 3));
        u1 = unit(pick(w, (dimension + 1) & 3));
        return;
    }
    const uint32_t d = camera ? dimension : dimension - 14;
    u0 = unit(owenScramble(sobol(uint32_t(id.index), d), pick(w, d & 3)));
    u1 = unit(owenScramble(sobol(uint32_t(id.index), d + 1), pick(w, (d + 1) & 3)));
}
// Two draws of an even dimension and the next (0/1 camera, 16/17 first-hit BSDF, 18/19 first-hit light), the Philox
// block they share computed once: Philox words of one group, or the Sobol pair's one seed block. The same words as
// two draw() calls.
__device__ __forceinline__ void drawPair(const Identity& id, uint32_t dimension, uint32_t policy, float& u0, float& u1) {
    const bool camera = dimension < 5, diffuse = dimension == 16, light = dimension == 18;
    const bool sobolPair = (camera && (policy & 1)) || (diffuse && (policy & 2)) || (light && (policy & 4));
    if (!sobolPair) {
        const Words w = block(id, dimension >> 2);
        u0 = unit(pick(w, dimension & 3));
        u1 = unit(pick(w, (dimension + 1) & 3));
        return;
    }
    const uint32_t d = camera ? dimension : dimension - 14;  // 0, 2 or 4: the pair (d, d + 1) within one seed block
    Identity shift = id;
    shift.index = id.index >> 32;
    const Words seeds = block(shift, 0x534f424fu + (d >> 2));
    u0 = unit(owenScramble(sobol(uint32_t(id.index), d), pick(seeds, d & 3)));
    u1 = unit(owenScramble(sobol(uint32_t(id.index), d + 1), pick(seeds, (d + 1) & 3)));
}

constexpr uint32_t CameraDimensions = 16, DimensionsPerBounce = 256;  // sample_stream.h:149

// A vertex's draws (async_integrator.h:151-176): dimension 16 + bounce*256 + offset on, one Philox block per four.
struct Stream {
    Identity id;
    uint32_t dimension, group;
    Words w;
    __device__ __forceinline__ Stream() : id{0, 0, 0, 0}, dimension(0), group(~0u), w{{0, 0, 0, 0}} {}
    __device__ __forceinline__ Stream(const Identity& i, uint32_t bounce)
        : id(i), dimension(CameraDimensions + bounce * DimensionsPerBounce), group(~0u), w{{0, 0, 0, 0}} {}
    // The group whose block the next draw needs, and that block computed elsewhere (block2) handed in.
    __device__ __forceinline__ uint32_t nextGroup() const { return dimension >> 2; }
    __device__ __forceinline__ void prime(const Words& b) {
        group = dimension >> 2;
        w = b;
    }
    __device__ __forceinline__ float next() {
        if ((dimension >> 2) != group) {
            group = dimension >> 2;
            w = block(id, group);
        }
        return unit(pick(w, dimension++ & 3));
    }
};

// A child's address in the sample's tree (async_integrator.h:281-286): a splitmix-style mix of parent and (a, b).
__device__ __forceinline__ uint64_t childBranch(uint64_t parent, uint32_t a, uint32_t b) {
    uint64_t z = parent + 0x9e3779b97f4a7c15ull * (uint64_t(a) * 0x10001u + uint64_t(b) + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace x2
