"""Write the teste_pedro golden fixture from the reference's data files.

Inputs (data only, no code): pedro/teste_pedro.json (configuration),
pedro/bloquinho_pedro/{rho,cp,cs}_aco.npy (700x550 material maps) and
pedro/results/result_teste_aco.npy (10000x64 float32 B-scan written by the
reference's Taichi 2D elastic backend for that configuration).
Output: tests/golden/pedro_teste_aco.npz (maps, B-scan, config text).
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pedro_teste_aco.npz")

with open(os.path.join(REF, "pedro", "teste_pedro.json")) as f:
    cfg_text = f.read()
maps = {k: np.load(os.path.join(REF, "pedro", "bloquinho_pedro", f"{k}_aco.npy"), allow_pickle=False)
        for k in ("rho", "cp", "cs")}
bscan = np.load(os.path.join(REF, "pedro", "results", "result_teste_aco.npy"), allow_pickle=False)
np.savez_compressed(OUT, rho=maps["rho"], cp=maps["cp"], cs=maps["cs"], bscan=bscan,
                    config=np.array(cfg_text))
print(OUT, os.path.getsize(OUT))
