"""Writes the small golden vectors under tests/golden/ from the oracle (run once; outputs are committed)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

if __name__ == "__main__":
    m = O.make_mesh(2, (4, 4), (0, 0), (0.9142, 1.0))
    np.savez(os.path.join(HERE, "layout_seed13.npz"), coords=O.layout_random(m, 3, 13))
