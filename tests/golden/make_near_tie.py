"""Writes tests/golden/near_tie_mixed3975.npz: a world state on a near-tie of the narrowphase.

The state is mixed world 3975 of the 4096-world batch (seed 1, 458 bodies, built by
pypybox2d_amd.builder.batch_from_counts) after 410 steps of the FP32 GPU pipeline as it
was when that pipeline still stored shape geometry in FP32.  In it, contact (fixture 229,
fixture 290) sits on the touching threshold to within the FP32 rounding of its shapes:
evaluated from the exact shapes it has one manifold point and touches, evaluated from
FP32-rounded shapes it has none, and the one-step positions of its island move by up to
2.5e-4.  The tests use it to show that the pipelines read exact geometry.

Run on a GPU host with an engine that rounds geometry to FP32 (the dump step), then here:
    python tests/golden/make_near_tie.py DUMP.npz
where DUMP.npz holds the 410-step state of the batch worlds listed in the dump (the
world whose index is 3975 is extracted).  The dump is made by
    e = Engine("fp32"); e.upload(batch); e.step_n(410, graph=True); st = e.download()
followed by L.concat of the chosen worlds; only the extraction is in this script.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pypybox2d_amd import layout as L  # noqa: E402

WORLD, INDEX_IN_DUMP = 3975, 15
OUT = os.path.join(HERE, "near_tie_mixed3975.npz")


def main(dump):
    d = np.load(dump)
    st = L.BatchArrays(*[d[k] for k in L._FIELDS])
    w = st.world(INDEX_IN_DUMP)
    assert w.body_d.shape[0] == 458, "the dump does not hold world %d at index %d" % (WORLD, INDEX_IN_DUMP)
    np.savez_compressed(OUT, **{k: getattr(w, k) for k in L._FIELDS})
    print("wrote", OUT)


if __name__ == "__main__":
    main(sys.argv[1])
