"""Seed initialisation of the canvas (processing/src/canvas.rs:73-84:
`SmallRng::seed_from_u64(0xdeadbeefdeadbeef).fill(&mut seeds)`, rand 0.8),
pinned by the published known answers of its parts (no GPU).

rand 0.8 is not vendored, and no reference test holds seed values, so the
composition stays "parity unpinned": which `seed_from_u64` expansion rand
0.8's SmallRng uses is read from rand_core 0.6's documented default (an LCG
advance plus the PCG32 XSH-RR output per 4 bytes; SmallRng does not override
it), and mode 2 offers the SplitMix64 seeding of rand_xoshiro's own
Xoshiro256PlusPlus::seed_from_u64. What is pinned here:
  * xoshiro256++ against its reference implementation's sequence from the
    state {1, 2, 3, 4} (the vector rand_xoshiro's tests carry),
  * SplitMix64 against its reference implementation's outputs for the seeds
    1234567 and 0,
  * the PCG32 XSH-RR output function against the pcg32 demo sequence of the
    PCG reference implementation (initstate 42, initseq 54),
and the library's fill (ht_seed_fill) against those pinned parts composed
with rand's byte-fill rule (8 bytes per next_u64, little-endian; a trailing
4-byte word is next_u32 = the high half)."""
import numpy as np

from hypertrace_amd.processing import REFERENCE_SEED, seed_fill

M64 = (1 << 64) - 1


def rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M64


def xoshiro256pp(s, n):
    s = list(s)
    out = []
    for _ in range(n):
        out.append((rotl((s[0] + s[3]) & M64, 23) + s[0]) & M64)
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
    return out


def splitmix64(state, n):
    out = []
    for _ in range(n):
        state = (state + 0x9E3779B97F4A7C15) & M64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        out.append(z ^ (z >> 31))
    return out


def xsh_rr(state):
    """PCG32's output function of a 64-bit state."""
    xs = (((state >> 18) ^ state) >> 27) & 0xFFFFFFFF
    rot = state >> 59
    return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF


PCG_MUL = 6364136223846793005


def pcg32_demo(initstate, initseq, n):
    """pcg32_srandom_r then pcg32_random_r of the PCG reference (output of
    the state before each advance)."""
    inc = ((initseq << 1) | 1) & M64
    state = 0

    def step():
        nonlocal state
        old = state
        state = (old * PCG_MUL + inc) & M64
        return xsh_rr(old)

    step()
    state = (state + initstate) & M64
    step()
    return [step() for _ in range(n)]


def rand_core_seed_words(seed, n=8):
    """rand_core 0.6's default seed_from_u64: advance first, then output."""
    inc = 11634580027462260723
    out = []
    for _ in range(n):
        seed = (seed * PCG_MUL + inc) & M64
        out.append(xsh_rr(seed))
    return out


def fill_u32(state, n):
    words = []
    outs = xoshiro256pp(state, (n + 1) // 2)
    for k in range(n // 2):
        words += [outs[k] & 0xFFFFFFFF, outs[k] >> 32]
    if n % 2:
        words.append(outs[-1] >> 32)
    return np.array(words, np.uint32)


def test_known_answers_of_the_parts():
    assert xoshiro256pp([1, 2, 3, 4], 10) == [
        41943041, 58720359, 3588806011781223, 3591011842654386, 9228616714210784205, 9973669472204895162,
        14011001112246962877, 12406186145184390807, 15849039046786891736, 10450023813501588000]
    assert splitmix64(1234567, 5) == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                      4593380528125082431, 16408922859458223821]
    assert splitmix64(0, 1) == [0xE220A8397B1DCDAF]
    assert pcg32_demo(42, 54, 6) == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_mode2_is_splitmix_seeded_xoshiro_fill():
    for seed in (1234567, REFERENCE_SEED, 0):
        for n in (1, 2, 7, 64):
            np.testing.assert_array_equal(seed_fill(n, 2, seed), fill_u32(splitmix64(seed, 4), n))


def test_mode0_is_the_rand_core_default_expansion():
    for seed in (REFERENCE_SEED, 1, 0x48595054):
        w = rand_core_seed_words(seed)
        state = [w[2 * i] | (w[2 * i + 1] << 32) for i in range(4)]
        for n in (1, 3, 10, 257):
            np.testing.assert_array_equal(seed_fill(n, 0, seed), fill_u32(state, n))
    # the two expansions differ: mode 0 is not mode 2
    assert not np.array_equal(seed_fill(16, 0, REFERENCE_SEED), seed_fill(16, 2, REFERENCE_SEED))


def test_fill_is_a_prefix_stream():
    # filling n words is the first n words of filling more (odd n ends on a
    # high half, so only even prefixes are shared)
    big = seed_fill(1000, 0, REFERENCE_SEED)
    for n in (2, 10, 998):
        np.testing.assert_array_equal(seed_fill(n, 0, REFERENCE_SEED), big[:n])
