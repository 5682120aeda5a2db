"""Write the golden fixtures under tests/golden from the reference's own data.

Run in a container where the reference is mounted (reads /root/reference):
    python tests/golden/make_fixtures.py
Fixtures:
  cart2sph_ref.npz   the reference's vendored cart->sph blocks l = 0..6
                     (dftax/basis/data/cart2sph.npz), as data
"""
from pathlib import Path

import numpy as np

REF = Path("/root/reference/dftax/basis/data/cart2sph.npz")
OUT = Path(__file__).resolve().parent


def main() -> None:
    with np.load(REF) as z:
        np.savez(OUT / "cart2sph_ref.npz", **{k: np.asarray(z[k]) for k in z.files})
    print("wrote", OUT / "cart2sph_ref.npz")


if __name__ == "__main__":
    main()
