"""Statistical image comparison: an estimate from R independent renders against a high-spp
reference image accumulated in FP64.

For every pixel channel, with v = sample variance across the R renders (each of spp samples):
  se^2(mean of the renders)  = v / R
  se^2(reference)            = v * spp / ref_spp   (the same per-sample variance, ref_spp samples)
  3-sigma test: |mean - ref| <= 3 sqrt(se^2 + se_ref^2)  (+ a float floor for noiseless pixels)
  relMSE = mean((mean - ref)^2 / (ref^2 + 0.01)), next to the value noise alone explains,
  mean((se^2 + se_ref^2) / (ref^2 + 0.01)).
An unbiased implementation of the same estimator passes the 3-sigma test on ~99.7% of Gaussian
pixel channels (path-tracing noise is heavier-tailed and the variance is itself estimated from R
renders, so the bar is 97%; two oracle estimates score ~98.3%) and has a relMSE close to the
noise-only value; a biased one (a lost term, a wrong weight, a 3% gain error) moves whole regions
of the image by many standard errors (~66% within).
"""
import numpy as np

EPS = 1e-2
WITHIN_MIN = 0.97
EXCESS_MAX = 2.0


def compare(renders, spp, ref, ref_spp):
    x = np.asarray(renders, np.float64)
    r = np.asarray(ref, np.float64)
    R = x.shape[0]
    mean = x.mean(0)
    v = x.var(0, ddof=1)
    se2 = v / R + v * spp / ref_spp
    d = mean - r
    tol = 3.0 * np.sqrt(se2) + 1e-6 * np.abs(r) + 1e-7
    within = float(np.mean(np.abs(d) <= tol))
    relmse = float(np.mean(d * d / (r * r + EPS)))
    noise = float(np.mean(se2 / (r * r + EPS)))
    return {"within_3sigma": within, "relmse": relmse, "relmse_noise": noise,
            "excess": relmse / max(noise, 1e-30)}


def passes(st) -> bool:
    return st["within_3sigma"] >= WITHIN_MIN and st["excess"] <= EXCESS_MAX
