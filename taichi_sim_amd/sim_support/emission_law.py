"""``.law`` delay-law files: per-shot emission delays and amplitudes.

Format (sim_support/emission_law.py:24-113): three header lines, then one
tab-separated row per (shot, element): numR numS numT numL numV retE ampE retR
ampR, where numT is the shot (numS the shot of a full-matrix capture) and numV
the 1-based element.  ``read_law`` returns (delay, amplitude), each
(shots x elements) float64.
"""
import numpy as np

_HEADER = ["# LOIS DE RETARD \n", "Version 1.0 \n",
           "numR\tnumS\tnumT\tnumL\tnumV\tretE\tampE\tretR\tampR\n"]
_COL = {"numR": 0, "numS": 1, "numT": 2, "numL": 3, "numV": 4,
        "retE": 5, "ampE": 6, "retR": 7, "ampR": 8}


class EmissionLaw:
    def write_law(self, root, emitter_law, receiver_law=None, emitter_amp=None, reciever_amp=None,
                  elem_range=None):
        emitter_law = np.asarray(emitter_law)
        if elem_range is None:
            elem_range = [0, emitter_law.shape[1] - 1]
        if receiver_law is None:
            receiver_law = emitter_law
        with open(root + ".law", "w") as f:
            f.writelines(_HEADER)
            for shot in range(emitter_law.shape[0]):
                for e in range(elem_range[0], elem_range[-1] + 1):
                    amp_e = emitter_amp[shot, e] if emitter_amp is not None else 1
                    amp_r = reciever_amp[shot, e] if reciever_amp is not None else 1
                    row = [0, 0, shot, 0, e + 1, emitter_law[shot, e], amp_e, receiver_law[shot, e], amp_r]
                    f.write("\t".join(f"{v}" for v in row) + "\n")

    def read_law(self, file_root):
        with open(file_root) as f:
            lines = f.readlines()
        data = np.array([ln[:-1].split('\t') for ln in lines[3:]], dtype=np.float32)
        n_elem = int(np.max(data[:, _COL["numV"]]))
        max_s = int(np.max(data[:, _COL["numS"]]))
        n_shots = (max_s + 1) if max_s != 0 else int(np.max(data[:, _COL["numT"]]) + 1)
        delay = np.zeros((n_shots, n_elem), dtype='float')
        amp = np.zeros((n_shots, n_elem), dtype='float')
        for s in range(n_shots):
            rows = data[s * n_elem:(s + 1) * n_elem]
            delay[s, :] = rows[:, _COL["retE"]]
            amp[s, :] = rows[:, _COL["ampE"]]
        return delay, amp
