"""Batched runs, host side (no GPU): the union mesh keeps every case's
cells, faces and slots in order and renumbers its boundary groups."""
import numpy as np

import axs
from test_gpu_batch import three_cases


def test_union_mesh_layout():
    cases = three_cases()
    b = axs.batch_case(cases)
    a = b.mesh.arrays
    off_c = off_f = 0
    for i, (c, (first, n)) in enumerate(zip(cases, b.info["ranges"])):
        m = c.mesh.arrays
        assert first == off_c and n == c.mesh.n_cells
        f = c.mesh.n_faces
        np.testing.assert_array_equal(a["face_owner"][off_f:off_f + f], m["face_owner"] + off_c)
        nb = m["face_neighbor"]
        np.testing.assert_array_equal(a["face_neighbor"][off_f:off_f + f], np.where(nb >= 0, nb + off_c, -1))
        g = m["face_group"]
        np.testing.assert_array_equal(a["face_group"][off_f:off_f + f], np.where(g >= 0, g + 4 * i, g))
        np.testing.assert_array_equal(a["cell_volume"][off_c:off_c + n], m["cell_volume"])
        off_c += n
        off_f += f
