"""Likelihood caller on the GPU path: the reference's pyGPUAAK class
(pygpuAAK/pygpuAAK.py:20-300) with the same constructor, data preparation and
methods, scoring many templates per call through the batched library entry
point instead of a Python loop.

Differences from the reference, each a fix of a defect there:
  * get_Fisher forms <d_i h, d_j h> over both channels; the reference takes
    <d_j h_I, d_j h_I> + <d_i h_II, d_j h_II> (pygpuAAK.py:293-295);
  * parameter conversion uses emri_kludge_suite_amd.convert (whole-name
    matching, D in Gpc), see that module;
  * the tdi and astropy imports of the reference are not needed.
"""
import numpy as np

from .convert import Converter
from .gpuAAK import GPUAAK
from .LISAnoise import LISA_Noise, generate_noise_frequencies, generate_noise_single_channel


class pyGPUAAK:
    def __init__(self, data_info, length, dt, init_dt, key_order, **kwargs):
        """data_info: {"channel1": .., "channel2": ..} (unweighted spectra) or
        {"injection_params": {key: value}} in the sampler's coordinates."""
        prop_defaults = {
            "T_fit": 1.0,
            "LISA": True,
            "backint": True,
            "eps": 1e-6,
            "nwalkers": 1,
            "ndevices": 1,
            "wd_dur": 1,
            "add_noise": None,
            "n_max": 0,
            "precision": "fp64",
            "seed": None,
        }
        self.ndim = len(key_order)
        self.length = int(length)
        self.init_dt = init_dt
        self.dt = dt
        self.converter = Converter("emri", key_order)
        self.key_order = list(key_order)
        self.total_time = self.length * self.dt
        self.init_length = int(np.ceil(self.total_time / self.init_dt))
        for prop, default in prop_defaults.items():
            setattr(self, prop, kwargs.get(prop, default))
        self.generator = GPUAAK(self.T_fit, self.init_length, self.length, self.init_dt, self.dt,
                                self.LISA, self.backint)
        self.generator.set_nmax(self.n_max)
        self.generator.set_precision(self.precision)
        self.rng = np.random.default_rng(self.seed)
        self.determine_freqs_noise()
        self.data_info = data_info
        self.create_input_data()

    # ------------------------------------------------------------ data
    def determine_freqs_noise(self):
        """rFFT grid with f0 = f1/10 and optional coloured noise (pygpuAAK.py:67-84)."""
        self.freqs = generate_noise_frequencies(self.length * self.dt, 1 / self.dt)
        self.fft_length = len(self.freqs)
        df = 1.0 / (self.length * self.dt)
        self.freqs[0] = self.freqs[1] / 10.0
        self.added_noise = [0.0, 0.0]
        if self.add_noise is not None:
            for k in range(2):
                self.added_noise[k] = generate_noise_single_channel(LISA_Noise, [], {"dur": self.wd_dur}, df,
                                                                    self.freqs, rng=self.rng)

    def create_input_data(self):
        """Weights sqrt(dt/N)/sqrt(S_n) and weighted data (pygpuAAK.py:86-145)."""
        w = 1.0 / np.sqrt(LISA_Noise(self.freqs, dur=self.wd_dur)) * np.sqrt(self.dt / self.length)
        self.noise_channel1_inv = w
        self.noise_channel2_inv = w.copy()
        if "channel1" in self.data_info:
            self.injection_channel1 = np.asarray(self.data_info["channel1"]).copy()
            self.injection_channel2 = np.asarray(self.data_info["channel2"]).copy()
        else:
            self.injection_params = self.data_info["injection_params"]
            zeros = np.zeros_like(w, dtype=np.complex128)
            ones = np.ones_like(w)
            self.generator.input_data(zeros, zeros, ones, ones)
            inj = np.array([[self.injection_params.get(key) for key in self.key_order]], dtype=np.float64).T
            hI_f, hII_f = self.getNLL(inj, return_TDI=True)
            self.injection_channel1_no_noise = hI_f
            self.injection_channel2_no_noise = hII_f
            self.injection_channel1 = hI_f[0] + self.added_noise[0]
            self.injection_channel2 = hII_f[0] + self.added_noise[1]
        self.data_channel1 = self.noise_channel1_inv * self.injection_channel1
        self.data_channel2 = self.noise_channel2_inv * self.injection_channel2
        self.generator.input_data(self.data_channel1, self.data_channel2, self.noise_channel1_inv,
                                  self.noise_channel2_inv)
        self.d_d = 4 * np.sum([np.abs(self.data_channel1) ** 2, np.abs(self.data_channel2) ** 2])

    # ------------------------------------------------------------ likelihood
    def NLL(self, iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D,
            return_snr=False, return_TDI=False, return_waveform=False, **kwargs):
        """1/2 (<d|d> + <h|h> - 2 <d|h>) per template (pygpuAAK.py:147-243)."""
        P = np.stack([np.atleast_1d(np.asarray(x, dtype=np.float64)) for x in (
            iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D)], axis=1)
        n = P.shape[0]
        if return_TDI or return_waveform:
            if return_TDI:
                hI = np.zeros((n, self.fft_length), dtype=np.complex128)
                hII = np.zeros((n, self.fft_length), dtype=np.complex128)
            else:
                hI = np.zeros((n, self.length + 2))
                hII = np.zeros((n, self.length + 2))
            for i in range(n):
                if return_TDI:
                    self.generator.WaveformThroughLikelihood(*P[i])
                    _, t1, t2 = self.generator.GetWaveform(is_Fourier=True)
                    hI[i].real, hI[i].imag = t1[0::2], t1[1::2]
                    hII[i].real, hII[i].imag = t2[0::2], t2[1::2]
                else:
                    self.generator.gpu_gen_AAK(*P[i])
                    _, hI[i], hII[i] = self.generator.GetWaveform(is_Fourier=False)
            return hI, hII
        like = self.generator.LikelihoodBatch(P)
        d_h, h_h = like[:, 0], like[:, 1]
        if return_snr:
            return np.sqrt(d_h), np.sqrt(h_h)
        return 1 / 2 * (self.d_d + h_h - 2 * d_h)

    def getNLL(self, x, **kwargs):
        """x: (ndim, n) in key order; periodic angles are wrapped in place
        (pygpuAAK.py:245-270)."""
        x = np.asarray(x, dtype=np.float64)
        if x.ndim == 1:
            x = x[:, None]
        self.converter.recycle(x)
        x_in = self.converter.physical_order(self.converter.convert(x.copy()))
        return self.NLL(*x_in, **kwargs)

    def get_Fisher(self, x):
        """Fisher matrix by central differences of the weighted spectra
        (pygpuAAK.py:272-300), with the inner product <d_i h, d_j h>."""
        x = np.asarray(x, dtype=np.float64)
        Mij = np.zeros((self.ndim, self.ndim))
        x_in = np.tile(x, (2 * self.ndim, 1))
        for i in range(self.ndim):
            x_in[2 * i, i] += self.eps
            x_in[2 * i + 1, i] -= self.eps
        hI, hII = self.getNLL(x_in.T, return_TDI=True)
        dI = [(hI[2 * i] - hI[2 * i + 1]) / (2 * self.eps) for i in range(self.ndim)]
        dII = [(hII[2 * i] - hII[2 * i + 1]) / (2 * self.eps) for i in range(self.ndim)]
        for i in range(self.ndim):
            for j in range(i, self.ndim):
                v = 4 * np.real(np.vdot(dI[i], dI[j]) + np.vdot(dII[i], dII[j]))
                Mij[i, j] = Mij[j, i] = v
        return Mij
