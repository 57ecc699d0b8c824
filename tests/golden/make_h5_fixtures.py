"""Writes the HDF5 fixtures of tests/test_cpu_hdf5.py with h5py, in the
calls the reference makes (field.py:144-151 for a PhysicalField,
field.py:718-736 for a VectorField, equation.py:362-372 for a trajectory),
and the arrays they hold as an .npz (no pickles).

Run with an interpreter that has h5py (this image: /opt/conda/bin/python3.9):
    /opt/conda/bin/python3.9 tests/golden/make_h5_fixtures.py
"""

import os

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main() -> None:
    rng = np.random.default_rng(2024)
    vec = rng.standard_normal((3, 5, 9, 5)) + 1j * rng.standard_normal((3, 5, 9, 5))
    phys = rng.standard_normal((8, 9, 8))
    traj = rng.standard_normal((3, 3, 5, 9, 5)) + 1j * rng.standard_normal((3, 3, 5, 9, 5))
    with h5py.File(os.path.join(HERE, "h5_vector_field.h5"), "w") as f:
        grp = f.create_group(str(12))
        grp.create_dataset("velocity_hat", data=vec, compression="gzip", compression_opts=9)
    with h5py.File(os.path.join(HERE, "h5_physical_field.h5"), "w") as f:
        grp = f.create_group(str(3))
        dset = grp.create_dataset("velocity_x", data=phys, compression="gzip", compression_opts=9)
        dset.attrs.create("time_step", 3)
    with h5py.File(os.path.join(HERE, "h5_trajectory.h5"), "w") as f:
        f.create_dataset("trajectory", data=traj, compression="gzip", compression_opts=9)
        chunks = f["trajectory"].chunks
    np.savez(os.path.join(HERE, "h5_fixtures_expected.npz"), vec=vec, phys=phys, traj=traj,
             traj_chunks=np.array(chunks))


if __name__ == "__main__":
    main()
