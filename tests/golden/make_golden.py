"""Writes tests/golden/kat.json: integer known-answer vectors for the random
number and Morton functions of the hybrid pass, computed by independent
pure-Python implementations that are first checked against published
vectors:
  - TEA (Wheeler & Needham 1994), 32 rounds, key 0, plaintext 0
    -> 41ea3a0a 94baa940; the reference's 16-round variant with its key
    constants (random.glsl:3-17) is the same routine with n = 16;
  - Numerical Recipes ranqd1 LCG (a = 1664525, c = 1013904223) from 0
    -> 3c6ef35f 47502932 d1ccf6e9 aaf95334 6252e503 9f2ec686 (random.glsl:21-27);
  - PCG32 XSH-RR (O'Neill 2014), pcg32_srandom_r(42, 54)
    -> a15c02b7 7b47f409 ba1d3330 83d2f293 bfa4784b cbed606e (pcg32-demo).
Run: python tests/golden/make_golden.py
"""
import json
import os

M32 = (1 << 32) - 1
M64 = (1 << 64) - 1


def tea_rounds(v0, v1, key, rounds, delta=0x9E3779B9):
    s = 0
    for _ in range(rounds):
        s = (s + delta) & M32
        v0 = (v0 + ((((v1 << 4) & M32) + key[0]) ^ (v1 + s) ^ ((v1 >> 5) + key[1]))) & M32
        v1 = (v1 + ((((v0 << 4) & M32) + key[2]) ^ (v0 + s) ^ ((v0 >> 5) + key[3]))) & M32
    return v0, v1


REF_KEY = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)


def tea16(a, b):
    return tea_rounds(a, b, REF_KEY, 16)[0]


def lcg_seq(state, n):
    out = []
    for _ in range(n):
        state = (1664525 * state + 1013904223) & M32
        out.append(state & 0x00FFFFFF)
    return out, state


class Pcg32:
    def __init__(self, initstate, initseq):
        self.state = 0
        self.inc = ((initseq << 1) | 1) & M64
        self.next()
        self.state = (self.state + initstate) & M64
        self.next()

    def next(self):
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & M64
        xs = (((old >> 18) ^ old) >> 27) & M32
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & M32


def morton3(x, y, z):
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


def check_published():
    assert tea_rounds(0, 0, (0, 0, 0, 0), 32) == (0x41EA3A0A, 0x94BAA940)
    state = 0
    seq = []
    for _ in range(6):
        state = (1664525 * state + 1013904223) & M32
        seq.append(state)
    assert seq == [0x3C6EF35F, 0x47502932, 0xD1CCF6E9, 0xAAF95334, 0x6252E503, 0x9F2EC686]
    g = Pcg32(42, 54)
    assert [g.next() for _ in range(6)] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def make():
    check_published()
    kat = {}
    pairs = [(0, 0), (1, 0), (0, 1), (12345, 0), (640 * 359 + 639, 0), (2073599, 7), (0xFFFFFFFF, 0xFFFFFFFF),
             (1920 * 540 + 960, 0xFFFFFFFF)]
    kat["tea16"] = [[a, b, tea16(a, b)] for a, b in pairs]
    kat["lcg"] = []
    for s in (0, 1, 0xDEADBEEF, tea16(5, 0)):
        vals, _ = lcg_seq(s, 8)
        kat["lcg"].append([s, vals])
    kat["rnd"] = [[s, [v / float(1 << 24) for v in lcg_seq(s, 4)[0]]] for s in (0, 7, tea16(100, 3))]
    kat["pcg32"] = []
    for st, sq in ((42, 54), (0, 0), (tea16(3, 0), (1 << 32) | 5), (0xFFFFFFFF, (3 << 32) | 123)):
        g = Pcg32(st, sq)
        kat["pcg32"].append([st, sq, [g.next() for _ in range(8)]])
    kat["u01"] = []
    for seed, purpose, index in ((tea16(0, 0), 1, 0), (tea16(77, 0), 2, 3), (tea16(9, 1), 3, 17)):
        g = Pcg32(seed, (purpose << 32) | index)
        kat["u01"].append([seed, purpose, index, [(g.next() >> 8) / float(1 << 24) for _ in range(4)]])
    kat["morton3"] = [[x, y, z, morton3(x, y, z)] for x, y, z in
                      ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1023, 1023, 1023), (512, 3, 700), (5, 1000, 64))]
    return kat


if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat.json")
    with open(out, "w") as f:
        json.dump(make(), f, indent=1)
    print("wrote", out)
