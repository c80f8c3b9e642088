#!/usr/bin/env python3
"""Generate the golden vectors of the likelihood (tests/golden/likelihood_golden.json).

The reference ships no stored likelihood value (its only CPU-vs-GPU check
prints differences, cuda/test/testLikelihoodMockup.cc:613-645), so the vectors
are computed here from the definitions, independently of the oracle's C code:
60-digit decimal arithmetic, plain Python, no numpy.

Definitions (doc/tex/likelihood.tex:9-55, partialWaveFit/pwaLikelihood.cc):
  element(I)_ij  = sum_ij / M                        (ampIntegralMatrix.cc:238-252)
  A_i(e)         = Abar_i(e) / sqrt(Re I^norm_ii)      (pwaLikelihood.cc:1253-1260)
  I^acc_ij       = I_ij / (sqrt(I^norm_ii) sqrt(I^norm_jj)), same-sector blocks
                                                      (pwaLikelihood.cc:1311-1334)
  sigma_e        = sum_r sum_s |sum_i T_rsi A_si(e)|^2 + f^2
  -lnL           = sum_e -ln sigma_e + nu (sum_rs sum_ij Re(T_rsi T*_rsj I^acc_ij) + f^2 eta)
                   - sum_rsi ln(1 / (1 + |T_rsi|^2 / gamma^2))    (half-Cauchy prior)
  dL/dRe T, dL/dIm T: Re, Im of sum_e (-2/sigma_e) S_rs(e) conj(A_si(e))
                   + 2 nu sum_j T_rsj conj(I^acc_ij) - prior'(|T|)/|T| * T
  dL/df          = sum_e (-2 f / sigma_e) + 2 nu f eta
with the parameter layout of buildParDataStruct (rank > sector (refl -1
first) > wave; T_rsi = 0 for i < r, real for i = r; flat last).  By default
nu = 1 (normalised amplitudes) and the acceptance integral is the
normalisation integral (eta = 1).  The acceptance cases integrate I^acc over an
accepted subset of N_acc of the M MC events (addAccIntegral,
pwaLikelihood.cc:1088-1107):
  n_acc          = accEventsOverride if given, else M (the normalisation
                   integral's event count)
  eta            = N_acc / n_acc
  element(I^acc) = sum over the accepted events / n_acc
then rescaled by sqrt(I^norm_ii I^norm_jj) as above.  The non-normalised case
(useNormalizedAmps(false)) keeps the raw amplitudes and I^acc, with
nu = N (the number of data events, pwaLikelihood.cc:253).

Inputs are drawn from a fixed numpy-free generator (Python's random.Random).
Run: python tests/golden/make_golden.py
"""
import json
import os
import random
from decimal import Decimal, getcontext

getcontext().prec = 60

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "likelihood_golden.json")

CASES = [
    # name, nW (refl -1, refl +1), rank, events, MC events, cauchy width (None: flat prior), seed
    ("w0+4_r1", (0, 4), 1, 30, 60, None, 1),
    ("w2+3_r2", (2, 3), 2, 40, 80, None, 2),
    ("w3+2_r1_cauchy", (3, 2), 1, 25, 50, 0.5, 3),
    ("w1+1_r1", (1, 1), 1, 20, 40, None, 4),
    ("w2+2_r2_cauchy", (2, 2), 2, 35, 70, 0.8, 5),
]
# acceptance and normalisation variants: name -> (fraction of MC events
# accepted, accEventsOverride, useNormalizedAmps)
VARIANTS = {}
CASES += [
    ("w2+3_r2_acc", (2, 3), 2, 40, 80, None, 6),
    ("w1+2_r1_acc_override_cauchy", (1, 2), 1, 30, 60, 0.5, 7),
    ("w2+2_r2_unnormalised_acc", (2, 2), 2, 30, 64, None, 8),
]
VARIANTS["w2+3_r2_acc"] = (0.6, 0, True)
VARIANTS["w1+2_r1_acc_override_cauchy"] = (0.45, 150, True)
VARIANTS["w2+2_r2_unnormalised_acc"] = (0.5, 0, False)


class C:
    """complex number over Decimal"""
    __slots__ = ("re", "im")

    def __init__(self, re, im=Decimal(0)):
        self.re = Decimal(re)
        self.im = Decimal(im)

    def __add__(self, o):
        return C(self.re + o.re, self.im + o.im)

    def __sub__(self, o):
        return C(self.re - o.re, self.im - o.im)

    def __mul__(self, o):
        if isinstance(o, C):
            return C(self.re * o.re - self.im * o.im, self.re * o.im + self.im * o.re)
        return C(self.re * o, self.im * o)

    def conj(self):
        return C(self.re, -self.im)

    def norm(self):
        return self.re * self.re + self.im * self.im


def draw_inputs(nW, rank, N, M, seed):
    rnd = random.Random(seed)
    W = nW[0] + nW[1]

    def amp():
        s = rnd.uniform(0.2, 1.0)
        return [rnd.gauss(0, 1) * s, rnd.gauss(0, 1) * s]

    scales = [rnd.uniform(0.3, 1.5) for _ in range(W)]
    data = [[[a * scales[i] for a in amp()] for i in range(W)] for _ in range(N)]
    mc = [[[a * scales[i] for a in amp()] for i in range(W)] for _ in range(M)]
    # parameters
    P = 0
    for r in range(rank):
        for s in range(2):
            n = nW[s] - r
            if n > 0:
                P += 2 * n - 1
    P += 1
    par = [rnd.uniform(-1.5, 1.5) for _ in range(P)]
    return data, mc, par, rnd


def golden(nW, rank, N, M, width, seed, variant=None):
    data, mc, par, rnd = draw_inputs(nW, rank, N, M, seed)
    accFrac, override, normalised = variant if variant else (None, 0, True)
    # accepted subset of the MC events (all of them without an acceptance variant)
    accIdx = [k for k in range(M) if rnd.random() < accFrac] if accFrac is not None else list(range(M))
    W = nW[0] + nW[1]
    gw = lambda s, j: j if s == 0 else nW[0] + j
    # integral sums over the MC events (full W x W)
    sums = [[C(0) for _ in range(W)] for _ in range(W)]
    for ev in mc:
        A = [C(Decimal(x[0]), Decimal(x[1])) for x in ev]
        for i in range(W):
            for j in range(W):
                sums[i][j] = sums[i][j] + A[i] * A[j].conj()
    accSums = [[C(0) for _ in range(W)] for _ in range(W)]
    for k in accIdx:
        A = [C(Decimal(x[0]), Decimal(x[1])) for x in mc[k]]
        for i in range(W):
            for j in range(W):
                accSums[i][j] = accSums[i][j] + A[i] * A[j].conj()
    Md = Decimal(M)
    elem = [[C(sums[i][j].re / Md, sums[i][j].im / Md) for j in range(W)] for i in range(W)]
    nAcc = Decimal(override if override else M)
    eta = Decimal(len(accIdx)) / nAcc
    accElem = [[C(accSums[i][j].re / nAcc, accSums[i][j].im / nAcc) for j in range(W)] for i in range(W)]
    sq = [elem[i][i].re.sqrt() if normalised else Decimal(1) for i in range(W)]
    nu = Decimal(1) if normalised else Decimal(N)
    # (normalised) data amplitudes and acceptance matrix (same-sector blocks)
    An = [[C(Decimal(ev[i][0]) / sq[i], Decimal(ev[i][1]) / sq[i]) for i in range(W)] for ev in data]
    Iacc = {}
    for s in range(2):
        for i in range(nW[s]):
            for j in range(nW[s]):
                a, b = gw(s, i), gw(s, j)
                Iacc[(s, i, j)] = C(accElem[a][b].re / (sq[a] * sq[b]), accElem[a][b].im / (sq[a] * sq[b]))
    # parameters -> production amplitudes
    T = {}
    idx = {}
    k = 0
    for r in range(rank):
        for s in range(2):
            for i in range(nW[s]):
                if i < r:
                    T[(r, s, i)] = C(0)
                elif i == r:
                    T[(r, s, i)] = C(Decimal(par[k]))
                    idx[(r, s, i)] = (k, None)
                    k += 1
                else:
                    T[(r, s, i)] = C(Decimal(par[k]), Decimal(par[k + 1]))
                    idx[(r, s, i)] = (k, k + 1)
                    k += 2
    f = Decimal(par[k])
    # event part
    lnl = Decimal(0)
    G = {key: C(0) for key in T}
    gflat = Decimal(0)
    for A in An:
        S = {}
        sigma = f * f
        for r in range(rank):
            for s in range(2):
                acc = C(0)
                for i in range(nW[s]):
                    acc = acc + T[(r, s, i)] * A[gw(s, i)]
                S[(r, s)] = acc
                sigma += acc.norm()
        lnl -= sigma.ln()
        w = Decimal(-2) / sigma
        for (r, s, i) in T:
            G[(r, s, i)] = G[(r, s, i)] + S[(r, s)] * A[gw(s, i)].conj() * w
        gflat += w * f
    # normalisation term: nu (sum Re(T_i T_j* I^acc_ij) + f^2 eta)
    norm = f * f * eta
    for r in range(rank):
        for s in range(2):
            for i in range(nW[s]):
                dn = C(0)
                for j in range(nW[s]):
                    I = Iacc[(s, i, j)]
                    norm += ((T[(r, s, i)] * T[(r, s, j)].conj()) * I).re
                    dn = dn + T[(r, s, j)] * I.conj()
                G[(r, s, i)] = G[(r, s, i)] + dn * (Decimal(2) * nu)
    gflat += Decimal(2) * nu * f * eta
    value = lnl + nu * norm
    # half-Cauchy prior
    if width is not None:
        g2 = Decimal(width) * Decimal(width)
        for key, t in T.items():
            r2 = t.norm()
            c = Decimal(1) / (Decimal(1) + r2 / g2)
            value -= c.ln()
            # (1/(r c)) * (-2 r g2 / (g2 + r2)^2) = -2 g2 / (c (g2 + r2)^2)
            factor = Decimal(-2) * g2 / (c * (g2 + r2) * (g2 + r2))
            G[key] = G[key] - t * factor
    grad = [None] * len(par)
    for key, (kr, ki) in idx.items():
        grad[kr] = float(G[key].re)
        if ki is not None:
            grad[ki] = float(G[key].im)
    grad[-1] = float(gflat)
    out = dict(nW=list(nW), rank=rank, events=N, mc_events=M, cauchy_width=width,
               data=data, mc=mc, par=par,
               integral_sums=[[[float(sums[i][j].re), float(sums[i][j].im)] for j in range(W)] for i in range(W)],
               value=float(value), event_part=float(lnl), norm_term=float(norm), gradient=grad)
    if variant:
        out.update(acc_index=accIdx, acc_events=len(accIdx), acc_override=override, normalised=normalised,
                   eta=float(eta),
                   acc_sums=[[[float(accSums[i][j].re), float(accSums[i][j].im)] for j in range(W)]
                             for i in range(W)])
    return out


def main():
    out = {"description": "golden vectors of -lnL, its gradient and the integral sums; made by "
                          "tests/golden/make_golden.py in 60-digit decimal arithmetic",
           "cases": {}}
    for name, nW, rank, N, M, width, seed in CASES:
        out["cases"][name] = golden(nW, rank, N, M, width, seed, VARIANTS.get(name))
        print(name, out["cases"][name]["value"])
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=0)
        fh.write("\n")


if __name__ == "__main__":
    main()
