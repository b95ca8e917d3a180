"""Copy the data the reference's SHiP template tests hold into tests/golden/ship/.

Runs in the development container only (it reads /root/reference/files).
Nothing is computed here and no reference code is run: the files are the
reference's own test inputs and outputs.

* materials/: the material json files the five template tests name
  (files/materials/{aisi1010,iron_pw,copper_water_cooled}.json), copied;
* <test>/parameters.csv: the template parameters (files/tests/<test>/parameters.csv);
* <test>/ref.npz: ``B_eval`` = B_eval_ref.csv (x, y, z, Bx, By, Bz at the
  probe points the test evaluates by nodal interpolation).  The field maps
  (map.npy) sit at the quadrature points of the reference's own gmsh mesh
  and are not copied.

    python tests/golden/make_ship_fixtures.py
"""
import os
import shutil

import numpy as np
import pandas as pd

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = '/root/reference/files'
OUT = os.path.join(HERE, 'ship')
TESTS = ['test_mag_1', 'test_mag_1_dil', 'test_mag_3', 'test_mag_3_dil', 'test_mag_4']
MATERIALS = ['aisi1010.json', 'iron_pw.json', 'copper_water_cooled.json']


def main():
    os.makedirs(os.path.join(OUT, 'materials'), exist_ok=True)
    for m in MATERIALS:
        shutil.copyfile(os.path.join(SRC, 'materials', m), os.path.join(OUT, 'materials', m))
    for t in TESTS:
        d = os.path.join(OUT, t)
        os.makedirs(d, exist_ok=True)
        shutil.copyfile(os.path.join(SRC, 'tests', t, 'parameters.csv'),
                        os.path.join(d, 'parameters.csv'))
        b_eval = pd.read_csv(os.path.join(SRC, 'tests', t, 'B_eval_ref.csv')).values
        np.savez_compressed(os.path.join(d, 'ref.npz'), B_eval=b_eval.astype(np.float64))
        print(t, b_eval.shape)


if __name__ == '__main__':
    main()
