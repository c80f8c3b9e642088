"""Learned initializer for batched fits: a multilayer perceptron (rootpwa_amd.mlp,
trained on the GPU by the library's own kernels) maps per-bin data-amplitude
moments to start values of the production amplitudes.  Every prediction is
refined by the plain fit and judged by the fit's lnL; random starts stay in the
batch as the fallback.

Moments of a bin, from the data amplitudes A normalised as the likelihood
normalises them, a_i = A_i / sqrt(I_ii) (partialWaveFit/pwaLikelihood.cc:
1253-1260), and scaled per event to unit norm (this removes any overall
per-event factor, whose heavy tail would otherwise dominate the noise):
    M_ij = mean_e a_i(e) conj(a_j(e)) / |a(e)|^2   (rpwa_data_moments_device, FP64 MFMA).
For a rank-1 intensity |T^T a|^2 the moments are, up to noise,
M ~ c (1 + conj(T) T^T / |T|^2): the dominant eigenvector of M is conj(T)/|T|.

Network input (z-scored with the training means and deviations): M_ii for all
waves; Re and Im of M_ij for nPairs off-diagonal pairs (every pair with wave 0,
the anchor whose phase the amplitudes are referred to, then the pairs with the
largest mean |M_ij| over the training bins); and the spectral estimate t, the
conjugated dominant eigenvector of M with t_0 real and positive (2 W - 1
reals).  Output (z-scored): the best fit's parameters (after the sign
convention) divided by sqrt(N), minus t (flat component zero): the network
learns a correction to the spectral estimate.  A start is decoded as
sqrt(N) (t + correction).  The global phase of T does not change the
likelihood."""
import time

import numpy as np

from . import batch as B
from . import capi, synth
from .mlp import MLP


def data_moments(d_data, nEvents, W, normSums, nMC, device=0):
    """M_ij of one bin (complex [W][W]) from device data amplitudes [nEvents][W]."""
    ptr = d_data.address if hasattr(d_data, "address") else d_data
    return capi.data_moments_device(ptr, nEvents, W, np.real(np.diag(normSums)) / float(nMC), True, device)


class Initializer:
    def __init__(self, nWaves, nPairs=256, nHidden=512, nHiddenLayers=3, seed=0, device=0):
        self.W = int(nWaves)
        self.P = 2 * self.W
        self.nPairs = int(nPairs)
        self.shape = (self.W + 2 * self.nPairs + 2 * self.W - 1, int(nHidden), int(nHiddenLayers), self.P)
        self.seed = int(seed)
        self.device = device
        self.pairs = None
        self.net = None
        self.lossHistory = None

    def _select_pairs(self, moments):
        iu = np.triu_indices(self.W, 1)
        mag = np.mean([np.abs(M[iu]) for M in moments], axis=0)
        anchor = iu[0] == 0
        # anchor pairs first, then by decreasing mean magnitude (stable)
        order = np.lexsort((-mag, ~anchor))[:self.nPairs]
        self.pairs = (iu[0][order], iu[1][order])

    def spectral(self, moments):
        """Spectral estimates [nBins][P] (parameter layout, flat 0): conj of the
        dominant eigenvector of M, anchor real and positive, unit norm."""
        W = self.W
        out = np.zeros((len(moments), 2 * W))
        for k, M in enumerate(moments):
            w, v = np.linalg.eigh(0.5 * (M + M.conj().T))
            t = np.conj(v[:, -1])
            t = t * np.exp(-1j * np.angle(t[0]))
            out[k, 0] = t[0].real
            out[k, 1:2 * W - 1:2] = t[1:].real
            out[k, 2:2 * W - 1:2] = t[1:].imag
        return out

    def features(self, moments, spec=None):
        spec = self.spectral(moments) if spec is None else spec
        rows = []
        for M, t in zip(moments, spec):
            sel = M[self.pairs]
            rows.append(np.concatenate([np.real(np.diag(M)), sel.real, sel.imag, t[:-1]]))
        return np.array(rows)

    def train(self, moments, bestParams, nEvents, epochs=400, batchSize=64, lr=1e-3, validation=0.1,
              checkEvery=10):
        """moments: list of [W][W]; bestParams: [nBins][P] best-fit parameters;
        nEvents: data counts.  A fraction `validation` of the bins (the last
        ones) is held out; training runs in blocks of checkEvery epochs and keeps
        the parameters with the lowest validation loss (early stopping)."""
        self._select_pairs(moments)
        spec = self.spectral(moments)
        X = self.features(moments, spec)
        T = np.asarray(bestParams, dtype=np.float64) / np.sqrt(np.asarray(nEvents, float))[:, None] - spec
        nVal = int(round(validation * len(X))) if validation > 0 else 0
        nTr = len(X) - nVal
        self.xMean, self.xStd = X[:nTr].mean(axis=0), X[:nTr].std(axis=0) + 1e-12
        self.tMean, self.tStd = T[:nTr].mean(axis=0), T[:nTr].std(axis=0) + 1e-12
        Xs, Ts = (X - self.xMean) / self.xStd, (T - self.tMean) / self.tStd
        self.net = MLP(*self.shape, seed=self.seed, device=self.device)
        hist, valHist = [], []
        best, bestParams_ = np.inf, None
        for e0 in range(0, epochs, checkEvery):
            n = min(checkEvery, epochs - e0)
            hist.extend(self.net.train(Xs[:nTr], Ts[:nTr], batchSize=min(batchSize, nTr), epochs=n, lr=lr,
                                       seed=self.seed + e0).tolist())
            if nVal:
                v, _ = self.net.lossGrad(Xs[nTr:], Ts[nTr:])
                valHist.append(v)
                if v < best:
                    best, bestParams_ = v, self.net.params()
        if bestParams_ is not None:
            self.net.setParams(bestParams_)
        self.lossHistory = np.array(hist)
        self.validationHistory = np.array(valHist)
        self.validationLoss = best
        return self.lossHistory

    def predict(self, moments, nEvents, spectralOnly=False):
        """Start values [nBins][P] for bins with these moments and data counts
        (spectralOnly: the spectral estimate without the network's correction)."""
        spec = self.spectral(moments)
        if spectralOnly:
            # a flat amplitude of exactly zero is a stationary point of the fit: start
            # it at the training bins' mean instead
            corr = np.zeros_like(spec)
            corr[:, -1] = self.tMean[-1] if self.net is not None else 0.01
        else:
            X = (self.features(moments, spec) - self.xMean) / self.xStd
            corr = self.net.forward(X).astype(np.float64) * self.tStd + self.tMean
        return (spec + corr) * np.sqrt(np.asarray(nEvents, dtype=np.float64))[:, None]


def fitted_bins(seed, sizes, W, nStarts, fitSeed, mcFactor=5, chunk=48, log=None, device=0):
    """Generates bins, fits each with nStarts random starts (batched) and returns
    per bin: moments, N, the best fit's parameters and value, all starts' values."""
    out = dict(moments=[], nEvents=[], best=[], bestValue=[], values=[], evals=[])
    for c0 in range(0, len(sizes), chunk):
        part = sizes[c0:c0 + chunk]
        bf = B.BatchFit(W, nStarts, device)
        for k, n in enumerate(part):
            d_data, sums, M, _ = synth.generate_bin(synth.bin_seed(seed, c0 + k), W, int(n), mcFactor, device)
            out["moments"].append(data_moments(d_data, int(n), W, sums, M, device))
            bf.addBinDevice(d_data, int(n), sums, M)
            d_data.free()
            out["nEvents"].append(int(n))
        bf.finish()
        par, val, st, ne = bf.fit(seed=fitSeed + c0)
        for k in range(len(part)):
            b = int(np.argmin(val[k]))
            out["best"].append(par[k, b])
            out["bestValue"].append(val[k, b])
            out["values"].append(val[k])
            out["evals"].append(ne[k])
        if log:
            log("fitted bins %d / %d" % (min(c0 + chunk, len(sizes)), len(sizes)))
        del bf
    for k in ("nEvents", "best", "bestValue", "values", "evals"):
        out[k] = np.array(out[k])
    return out


def starts_needed(values, best, tol=0.5):
    """1-based index of the first start within tol of best (len + 1 if none)."""
    ok = np.flatnonzero(values <= best + tol)
    return int(ok[0]) + 1 if ok.size else len(values) + 1


def evaluate(init, seed, sizes, W, nRandom=64, nMixed=8, fitSeed=77, mcFactor=5, chunk=48, log=None, device=0):
    """Held-out comparison, three arms per bin: nRandom random starts; the
    network's prediction plus the first nMixed - 1 of those random starts; the
    spectral estimate alone (the network's input, without the network) plus the
    same nMixed - 1 random starts.  Reports the starts and evaluations needed to
    come within 0.5 of the best lnL found by any arm, and the wall time of the
    batched fits."""
    rows = []
    tR = tM = tS = 0.0
    for c0 in range(0, len(sizes), chunk):
        part = sizes[c0:c0 + chunk]
        bfR = B.BatchFit(W, nRandom, device)
        bfM = B.BatchFit(W, nMixed, device)
        bfS = B.BatchFit(W, nMixed, device)
        moments, nev = [], []
        for k, n in enumerate(part):
            d_data, sums, M, _ = synth.generate_bin(synth.bin_seed(seed, c0 + k), W, int(n), mcFactor, device)
            moments.append(data_moments(d_data, int(n), W, sums, M, device))
            bfR.addBinDevice(d_data, int(n), sums, M)
            bfM.addBinDevice(d_data, int(n), sums, M)
            bfS.addBinDevice(d_data, int(n), sums, M)
            d_data.free()
            nev.append(int(n))
        bfR.finish()
        bfM.finish()
        bfS.finish()
        rnd = bfR.startValues(seed=fitSeed + c0)
        mixed = np.empty((len(part), nMixed, 2 * W))
        mixed[:, 0] = init.predict(moments, nev)
        mixed[:, 1:] = rnd[:, :nMixed - 1]
        startVal, _ = bfM.eval(mixed)  # -lnL at the start points (MLP, random)
        capi.synchronize()
        t = time.perf_counter()
        parR, valR, stR, neR = bfR.fit(startPar=rnd)
        tR += time.perf_counter() - t
        t = time.perf_counter()
        parM, valM, stM, neM = bfM.fit(startPar=mixed)
        tM += time.perf_counter() - t
        spectral = np.empty_like(mixed)
        spectral[:, 0] = init.predict(moments, nev, spectralOnly=True)
        spectral[:, 1:] = rnd[:, :nMixed - 1]
        startValS, _ = bfS.eval(spectral)  # -lnL at the spectral start (the selection arm compares it)
        t = time.perf_counter()
        parS, valS, stS, neS = bfS.fit(startPar=spectral)
        tS += time.perf_counter() - t
        for k in range(len(part)):
            best = min(valR[k].min(), valM[k].min(), valS[k].min())
            sR = starts_needed(valR[k], best)
            sM = starts_needed(valM[k], best)
            sS = starts_needed(valS[k], best)
            # selection arm: the network's start or the spectral one, whichever has the lower -lnL
            # (one more evaluation), then the same random starts
            useMlp = startVal[k][0] <= startValS[k][0]
            sSel = sM if useMlp else sS
            eSel = 1 + int((neM[k][:min(sM, nMixed)] if useMlp else neS[k][:min(sS, nMixed)]).sum())
            rows.append(dict(events=nev[k], best=best, startsRandom=sR, startsMixed=sM, startsSpectral=sS,
                             startsSelected=sSel, selectedMlp=bool(useMlp), evalsSelected=eSel,
                             selectedReaches=bool((valM[k][0] if useMlp else valS[k][0]) <= best + 0.5),
                             evalsRandom=int(neR[k][:min(sR, nRandom)].sum()),
                             evalsMixed=int(neM[k][:min(sM, nMixed)].sum()),
                             evalsSpectral=int(neS[k][:min(sS, nMixed)].sum()),
                             mlpValue=float(valM[k][0]), mlpEvals=int(neM[k][0]),
                             mlpStartGap=float(startVal[k][0] - best),
                             spectralStartGap=float(startValS[k][0] - best),
                             randomStartGap=float(np.median(startVal[k][1:] - best)),
                             randomEvalsMean=float(neR[k].mean()), randomWithin=int(np.sum(valR[k] <= best + 0.5)),
                             spectralReaches=bool(valS[k][0] <= best + 0.5), spectralEvals=int(neS[k][0])))
        if log:
            log("evaluated bins %d / %d" % (min(c0 + chunk, len(sizes)), len(sizes)))
        del bfR, bfM, bfS
    sR = np.array([r["startsRandom"] for r in rows])
    sM = np.array([r["startsMixed"] for r in rows])
    sS = np.array([r["startsSpectral"] for r in rows])
    eR = np.array([r["evalsRandom"] for r in rows])
    eM = np.array([r["evalsMixed"] for r in rows])
    eS = np.array([r["evalsSpectral"] for r in rows])
    sSel = np.array([r["startsSelected"] for r in rows])
    eSel = np.array([r["evalsSelected"] for r in rows])
    return {
        "bins": len(rows),
        "selection_arm": {"note": "per bin the network's or the spectral start, whichever has the lower -lnL "
                                  "(+1 evaluation), then the same random starts",
                          "network_chosen": float(np.mean([r["selectedMlp"] for r in rows])),
                          "starts_needed": float(sSel.mean()),
                          "first_start_reaches_best": float(np.mean([r["selectedReaches"] for r in rows])),
                          "evaluations_needed": float(eSel.mean())},
        "starts_needed": {"random": float(sR.mean()), "mlp_plus_random": float(sM.mean()),
                          "spectral_plus_random": float(sS.mean()), "ratio": float(sR.mean() / sM.mean())},
        "reached_within_budget": {"random_%d" % nRandom: float(np.mean(sR <= nRandom)),
                                  "mlp_plus_%d" % (nMixed - 1): float(np.mean(sM <= nMixed)),
                                  "spectral_plus_%d" % (nMixed - 1): float(np.mean(sS <= nMixed))},
        "mlp_start_reaches_best": float(np.mean(sM == 1)),
        "first_random_start_reaches_best": float(np.mean(sR == 1)),
        "start_gap_median": {"mlp": float(np.median([r["mlpStartGap"] for r in rows])),
                             "spectral": float(np.median([r["spectralStartGap"] for r in rows])),
                             "random": float(np.median([r["randomStartGap"] for r in rows]))},
        "evaluations_needed": {"random": float(eR.mean()), "mlp_plus_random": float(eM.mean()),
                               "spectral_plus_random": float(eS.mean()),
                               "ratio": float(eR.mean() / max(eM.mean(), 1e-9)),
                               "mlp_over_spectral": float(eM.mean() / max(eS.mean(), 1e-9))},
        "evaluations_per_fit": {"random_start": float(np.mean([r["randomEvalsMean"] for r in rows])),
                                "mlp_start": float(np.mean([r["mlpEvals"] for r in rows])),
                                "spectral_start_without_network": float(np.mean([r["spectralEvals"] for r in rows]))},
        "spectral_start_without_network_reaches_best": float(np.mean([r["spectralReaches"] for r in rows])),
        "wall_s": {"random_%d_starts" % nRandom: tR, "mlp_plus_%d" % (nMixed - 1): tM,
                   "spectral_plus_%d" % (nMixed - 1): tS, "ratio": tR / max(tM, 1e-9)},
        "per_bin": rows,
    }
