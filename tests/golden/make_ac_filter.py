"""Designs the default AC (low-pass) filter once and writes it as a
Filter_Mode=2 vector [order, isFIR, b...] (src/HFunction.cpp:137-153).

The reference designs this filter at run time with Parks-McClellan from its
default band spec (src/params.cpp:61-68: Fpass 600 Hz, Fstop 1600 Hz, Apass
3 dB, Astop 60 dB, equiripple, minimum order).  Here the same spec goes to
scipy.signal.remez; the order is the Herrmann/Kaiser minimum-order estimate,
the band weights follow the reference's rule (stop weight 10, pass weight
10*delta_stop/delta_pass, src/SubModel.cpp:224-232).
The fixtures it writes live in cochlearmodel_amd/data/, where the library and
the tests read them.
Run: python tests/golden/make_ac_filter.py
"""
import json
import math
import os

import numpy as np
from scipy import signal

HERE = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                    "cochlearmodel_amd", "data")


def design(fs=44100.0, fpass=600.0, fstop=1600.0, apass=3.0, astop=60.0):
    dp = (10 ** (apass / 20) - 1) / (10 ** (apass / 20) + 1)
    ds = 10 ** (-astop / 20)
    df = (fstop - fpass) / fs
    n = int(math.ceil((-20 * math.log10(math.sqrt(dp * ds)) - 13) / (14.6 * df))) + 1
    w_pass = max(0.01, 10 * ds / dp)
    w_stop = 10.0
    b = signal.remez(n, [0, fpass, fstop, fs / 2], [1, 0], weight=[w_pass, w_stop], fs=fs, maxiter=200)
    return b


if __name__ == "__main__":
    for fs in (44100.0, 20000.0):
        b = design(fs)
        vec = [float(len(b)), 1.0] + [float(x) for x in b]
        path = os.path.join(HERE, f"ac_filter_equiripple_{int(fs)}.json")
        with open(path, "w") as f:
            json.dump({"fs": fs, "spec": {"Fpass": 600, "Fstop": 1600, "Apass": 3, "Astop": 60},
                       "filter_mode_2": vec}, f, indent=0)
        w, h = signal.freqz(b, worN=4096, fs=fs)
        print(path, "taps", len(b), "gain@0", abs(h[0]), "max stop", 20 * np.log10(abs(h[w >= 1600]).max()))
