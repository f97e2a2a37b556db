"""Write tests/golden/dipole_grid.npz: the golden scattered-field coefficients
that the reference's own test holds for a dipole-illuminated grid of 21
spheres with l_max 4..7 in one medium, solved with the explicit coupling and
LU (tests/system_tests/test_direct_radial_and_volume_particle_coupling.py:
26-80, the '2D' and '3D' cases: 1748 complex coefficients each).

Only the two numeric lists assigned to `known_result` are read (as literals
from the file's syntax tree); the geometry is restated in the tests that use
them. Run in a container that has the reference checkout:
    python tests/golden/make_dipole_grid_fixture.py /root/reference
"""
import ast
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    path = os.path.join(ref_root, 'tests', 'system_tests', 'test_direct_radial_and_volume_particle_coupling.py')
    tree = ast.parse(open(path).read())
    lists = []
    for node in ast.walk(tree):
        if isinstance(node, ast.Assign) and any(isinstance(t, ast.Name) and t.id == 'known_result'
                                                for t in node.targets):
            if isinstance(node.value, ast.List) and len(node.value.elts) > 100:
                lists.append(np.array(ast.literal_eval(node.value), dtype=np.complex128))
    if len(lists) != 2:
        raise SystemExit('expected the 2D and 3D result lists, found %d' % len(lists))
    np.savez(os.path.join(HERE, 'dipole_grid.npz'), known_2d=lists[0], known_3d=lists[1])
    print('wrote', len(lists[0]), len(lists[1]), 'coefficients')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
