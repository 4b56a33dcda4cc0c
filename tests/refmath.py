"""Independent numpy formulations used to pin the oracle.

These are written from the mathematics (complex matrix products, array
transposes), not by retyping the oracle's loops, so that a dropped term, a
wrong sign or a transposed index in the oracle shows up as a mismatch.
"""
import numpy as np


def taftp_to_complex(taftp, nhg, A, F, S=256, P=2):
    v = taftp.reshape(nhg, A, F, S, P, 2).astype(np.int64)
    return v[..., 0] + 1j * v[..., 1]


def split_transpose_np(taftp, nhg, A_tot, A_used, a0, F, S=256, P=2):
    v = taftp.reshape(nhg, A_tot, F, S, P, 2)[:, a0:a0 + A_used]
    # -> [F][hg][S][P][A][2]
    return np.ascontiguousarray(v.transpose(2, 0, 3, 4, 1, 5)).reshape(-1)


def ftpa_complex(ftpa, F, T, A, P=2):
    v = ftpa.reshape(F, T, P, A, 2).astype(np.float64)
    return v[..., 0] + 1j * v[..., 1]


def fba_complex(fba, F, B, A):
    w = fba.reshape(F, B, A, 2).astype(np.float64)
    return w[..., 0] + 1j * w[..., 1]


def cb_sums_np(ftpa, fba, F, T, A, B, P=2):
    """beam[f,t,p,b] = sum_a w[f,b,a] v[f,t,p,a] (exact in float64 here: |terms| < 2^53)."""
    v = ftpa_complex(ftpa, F, T, A, P)           # F T P A
    w = fba_complex(fba, F, B, A)                # F B A
    beam = np.einsum("ftpa,fba->ftpb", v, w)     # no conjugate
    out = np.empty(beam.shape + (2,), dtype=np.int64)
    out[..., 0] = np.rint(beam.real).astype(np.int64)
    out[..., 1] = np.rint(beam.imag).astype(np.int64)
    return out


def cb_power_np(ftpa, fba, F, T, A, B, ts, fs, P=2):
    s = cb_sums_np(ftpa, fba, F, T, A, B, P)
    pw = (s[..., 0] ** 2 + s[..., 1] ** 2)       # F T P B  (int64 exact)
    pw = pw.sum(axis=2)                          # F T B
    pw = pw.reshape(F // fs, fs, T // ts, ts, B).sum(axis=(1, 3))
    return pw.astype(np.float64)                 # Fo To B


def tbtf_from_fo_to_b(q, ohs):
    """[Fo][To][B] int8 -> TBTF bytes [To/ohs][B][ohs][Fo]."""
    Fo, To, B = q.shape
    x = q.reshape(Fo, To // ohs, ohs, B).transpose(1, 3, 2, 0)
    return np.ascontiguousarray(x).reshape(-1)


def ib_power_np(taftp, nhg, A, F, ts, fs, S=256, P=2):
    v = taftp.reshape(nhg, A, F, S, P, 2).astype(np.int64)
    p = (v ** 2).sum(axis=(-1, -2, 1))           # hg F S
    p = p.reshape(nhg, F // fs, fs, S // ts, ts).sum(axis=(2, 4))   # hg Fo S/ts
    return p.transpose(0, 2, 1).reshape(nhg * (S // ts), F // fs).astype(np.float64)


def quantise_np(power, offset, scale, mode=0):
    v = (np.asarray(power, dtype=np.float64) - np.float64(np.float32(offset))) / np.float64(np.float32(scale))
    q = np.trunc(v)
    if mode == 0:
        frac = v - q                             # exact
        q = q + np.where(np.abs(frac) >= 0.5, np.sign(v), 0.0)
    q = np.where(np.isnan(q), -128, q)
    return np.clip(q, -128, 127).astype(np.int8)
