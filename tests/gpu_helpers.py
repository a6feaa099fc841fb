"""Shared helpers for the GPU parity tests (seeded inputs, oracle calls)."""
import numpy as np

from sarsolver_amd import synthetic


def geometry(ap):
    return (np.asarray(ap.transmitter_positions), np.asarray(ap.receiver_positions),
            np.asarray(ap.stab_ref_posns))


def oracle_forward(O, ap, pos, x, U, interp, threads=8):
    tx, rx, srp = geometry(ap)
    return O.forward(tx, rx, srp, pos, x.reshape(-1), np.fft.ifftshift(ap.waveform_fft),
                     ap.slow_time_weighting, ap.centre_frequency, ap.sample_bandwidth,
                     ap.num_freqs, c_eff=ap.c_eff, upsample_ratio=U,
                     sign_multiplier=ap.sign_multiplier, interp=interp, num_threads=threads)


def oracle_adjoint(O, ap, pos, y_fft, U, interp, threads=8):
    tx, rx, srp = geometry(ap)
    return O.adjoint(tx, rx, srp, pos, y_fft, np.fft.ifftshift(ap.waveform_fft),
                     ap.slow_time_weighting, ap.centre_frequency, ap.sample_bandwidth,
                     c_eff=ap.c_eff, upsample_ratio=U, sign_multiplier=ap.sign_multiplier,
                     interp=interp, num_threads=threads)


def plan(ap, sc_pos, U, interp, precision="fp64", dev=None):
    from sarsolver_amd import SarPlan
    tx, rx, srp = geometry(ap)
    return SarPlan(tx, rx, srp, sc_pos, np.fft.ifftshift(ap.waveform_fft), ap.slow_time_weighting,
                   ap.centre_frequency, ap.sample_bandwidth, ap.c_eff, U, ap.sign_multiplier,
                   interp=interp, precision=precision, device=dev)


def c1_inputs(seed_offset=0):
    w = synthetic.WORKLOADS["c1_small"]
    rng = np.random.default_rng(w.seed + seed_offset)
    ap, sc = synthetic.build(w)
    x = synthetic.scene_amplitudes(w, rng).reshape(-1)
    y = synthetic.complex_normal(rng, (w.pulses, w.samples))
    return w, ap, sc, x, y
