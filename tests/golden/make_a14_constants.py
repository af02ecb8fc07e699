"""Writes tests/golden/a14_constants.json from the reference's A14 dam-break pack (data only: constants and a
sample of its initial positions). Reads bench/xbench/ours/packs/laws/A14.dam_break/a50_r1mm/{fluid.json,
particles.npz} under /root/reference with json and numpy.load(allow_pickle=False)."""

import json
import os
import re

import numpy as np

SRC = "/root/reference/bench/xbench/ours/packs/laws/A14.dam_break/a50_r1mm"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "a14_constants.json")

with open(os.path.join(SRC, "fluid.json")) as f:
    d = json.load(f)
sph = d["sph"]
q = np.load(os.path.join(SRC, "particles.npz"), allow_pickle=False)["particle_q"]
out = {
    "source": "A14.dam_break/a50_r1mm/fluid.json and particles.npz of the reference",
    "particle_radius": sph["particle_radius"]["value"],
    "spacing": sph["spacing"]["value"],
    "support_radius": sph["support_radius"]["value"],
    "kernel": sph["kernel"]["value"],
    "rest_density": sph["rest_density"]["value"],
    "particle_mass": sph["particle_mass"]["value"],
    "lattice_sum_S": float(re.search(r"S = (\d+)", sph["particle_mass"]["meaning"]).group(1)),
    "lattice_sum_S_note": "parsed from fluid.json sph.particle_mass.meaning (m = rho0 / S)",
    "lattice_neighbours": sph["lattice_neighbours"]["value"],
    "n_particles": int(q.shape[0]),
    "first_positions": q[:16].astype(float).tolist(),
    "bbox_lo": q.min(0).astype(float).tolist(),
    "bbox_hi": q.max(0).astype(float).tolist(),
}
with open(OUT, "w") as f:
    json.dump(out, f, indent=1)
print("wrote", OUT)
