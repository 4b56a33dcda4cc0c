"""Writes tests/golden/cb_small.json: a small coherent-beam case whose expected
int32 sums and detected power are computed here by plain Python integer loops
straight from the definition (CoherentBeamformerTester.cu:55-100):

    beam(f,t,p,b) = sum_a (wr*vr - wi*vi) + i (wr*vi + wi*vr)
    power(fo,to,b) = sum over fscrunch, tscrunch, pol of re^2 + im^2

No numpy arithmetic and no oracle code is involved, so the fixture pins the
oracle independently.  Inputs come from Python's own `random` with a fixed
seed; the committed JSON, not this script, is the fixture.
"""
import cmath
import json
import math
import os
import random

F, T, A, B, TS, FS, P = 2, 8, 4, 8, 4, 2, 2


def main():
    r = random.Random(20240601)
    ftpa = [max(-127, min(127, round(r.gauss(0.0, 16.0)))) for _ in range(F * T * P * A * 2)]
    weights = []
    for _ in range(F * B * A):
        th = r.uniform(0.0, 2 * math.pi)
        z = 127 * cmath.exp(1j * th)
        weights += [int(math.floor(z.real + 0.5)) if z.real >= 0 else -int(math.floor(-z.real + 0.5)),
                    int(math.floor(z.imag + 0.5)) if z.imag >= 0 else -int(math.floor(-z.imag + 0.5))]
    sums = []
    for f in range(F):
        for t in range(T):
            for p in range(P):
                for b in range(B):
                    re = im = 0
                    for a in range(A):
                        vi = ((f * T + t) * P + p) * A + a
                        wi = (f * B + b) * A + a
                        vr_, vi_ = ftpa[2 * vi], ftpa[2 * vi + 1]
                        wr_, wi_ = weights[2 * wi], weights[2 * wi + 1]
                        re += wr_ * vr_ - wi_ * vi_
                        im += wr_ * vi_ + wi_ * vr_
                    sums += [re, im]
    power = []
    for fo in range(F // FS):
        for to in range(T // TS):
            for b in range(B):
                acc = 0
                for fc in range(FS):
                    for ts in range(TS):
                        for p in range(P):
                            f, t = fo * FS + fc, to * TS + ts
                            idx = (((f * T + t) * P + p) * B + b) * 2
                            acc += sums[idx] ** 2 + sums[idx + 1] ** 2
                power.append(acc)
    out = dict(F=F, T=T, A=A, B=B, ts=TS, fs=FS, ftpa=ftpa, weights=weights, sums=sums, power=power)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cb_small.json")
    with open(path, "w") as fh:
        json.dump(out, fh)


if __name__ == "__main__":
    main()
