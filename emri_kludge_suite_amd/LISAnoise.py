"""Analytic LISA noise model and rFFT frequency grid (pygpuAAK/LISAnoise.py).

Same formulas as the reference: instrument noise (OMS + acceleration, with the
f/f_star transfer factor) plus the galactic confusion foreground S_c of the
given observation duration (0 or None disables it).
"""
import numpy as np

# galactic confusion fit parameters per observation time in years (LISAnoise.py:3-8)
FIT_PARS = {
    0.5: {"alpha": 0.133, "beta": 243, "kappa": 482, "gamma": 917, "f_k": 2.58e-3},
    1: {"alpha": 0.171, "beta": 292, "kappa": 1020, "gamma": 1680, "f_k": 2.15e-3},
    2: {"alpha": 0.165, "beta": 299, "kappa": 611, "gamma": 1340, "f_k": 1.73e-3},
    4: {"alpha": 0.138, "beta": -221, "kappa": 521, "gamma": 1680, "f_k": 1.13e-3},
}


def P_OMS(f):
    return (1.5e-11) ** 2 * (1 + (2e-3 / f) ** 4)


def P_acc(f):
    return (3e-15) ** 2 * (1 + (0.4e-3 / f) ** 2) * (1 + (f / 8e-3) ** 4)


def S_c(f, dur=4):
    if dur == 0 or dur is None:
        return 0.0
    if dur not in FIT_PARS:
        raise ValueError("dur needs to be 0.5, 1, 2, or 4 years.")
    fp = FIT_PARS[dur]
    return (9e-45 * f ** (-7 / 3) * np.exp(-(f ** fp["alpha"]) + fp["beta"] * f * np.sin(fp["kappa"] * f))
            * (1 + np.tanh(fp["gamma"] * (fp["f_k"] - f))))


def LISA_Noise(f, L=2.5e9, f_star=19.09e-3, dur=4):
    """One-sided PSD S_n(f) (LISAnoise.py:39-43)."""
    return (10.0 / (3.0 * L ** 2) * (P_OMS(f) + 2 * (1 + np.cos(f / f_star) ** 2) * P_acc(f) / (2 * np.pi * f) ** 4)
            * (1 + 6 / 10 * (f / f_star) ** 2) + S_c(f, dur=dur))


def generate_noise_frequencies(Tobs, fs):
    """rFFT frequency grid 0 .. fs/2 (LISAnoise.py:46-53)."""
    n = int(np.round(Tobs * fs))
    nf = int(np.round(n / 2) + 1)
    return np.linspace(start=0, stop=fs / 2, num=nf)


def generate_noise_single_channel(noise_func, noise_args, noise_kwargs, df, data_freqs, rng=None):
    """Coloured Gaussian noise with the reference normalisation 0.5/sqrt(df)
    per real/imaginary part (LISAnoise.py:56-63)."""
    rng = np.random.default_rng() if rng is None else rng
    norm1 = 0.5 * (1.0 / df) ** 0.5
    htilde = rng.normal(0, norm1, data_freqs.shape) + 1j * rng.normal(0, norm1, data_freqs.shape)
    return np.sqrt(noise_func(data_freqs, *noise_args, **noise_kwargs)) * htilde
