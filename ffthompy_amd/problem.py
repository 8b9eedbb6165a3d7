"""Problem definition and input files (ffthompy/problem.py:9-160).

An input file is a Python module exporting `materials` (dict) and `problems` (list of
dicts with 'name', 'physics', 'material', 'solve', 'solver', 'postprocess' and an
optional 'save'), as documented in the reference's users' guide.
"""
import os
import pickle
import sys

import numpy as np

from . import applications


class Problem:
    """Parse one problem of an input file, run it, post-process it (problem.py:9-140)."""

    def __init__(self, conf_problem=None, conf=None):
        self.__dict__.update(conf_problem)
        if isinstance(self.material, str):
            conf_material = conf.materials[self.material]
        else:
            conf_material = self.material
        self.material = self.parse_material(conf_material)
        self.Y = np.array(self.material['Y'], dtype=np.float64)
        self.dim = self.Y.size
        if self.physics == 'scalar':
            self.shape = (self.dim,)
        elif self.physics == 'elasticity':
            self.shape = (int(self.dim * (self.dim + 1) / 2),)
        self.output = {}

    @staticmethod
    def parse_material(conf_material):
        """'all' / 'otherwise' may appear once and last (problem.py:43-68)."""
        if 'fun' in conf_material:
            return conf_material
        incl = conf_material['inclusions']
        for key in ('all', 'otherwise'):
            n = incl.count(key)
            if n > 1:
                raise ValueError("Maximal one occurrence of inclusion 'otherwise' or 'all' is allowed!")
            if n == 1 and incl.index(key) != len(incl) - 1:
                raise ValueError('inclusion ("all" or "otherwise") has to be at the end of the list')
        return conf_material

    def calculate(self):
        """Dispatch on physics (problem.py:70-83)."""
        fun = getattr(applications, self.physics, None)
        if fun is None:
            raise NotImplementedError('Not implemented physics (%s).' % self.physics)
        fun(self)

    def postprocessing(self):
        """Print the homogenized matrices; save host copies of the output when asked
        (problem.py:85-126)."""
        if self.physics in ('scalar', 'elasticity'):
            print('\nHomogenized matrices')
            for pd in self.solve['primaldual']:
                for key, val in self.output['mat_' + pd].items():
                    print(key)
                    print(val)
        if hasattr(self, 'save'):
            filename = self.save['filename']
            dirs = os.path.dirname(filename)
            if dirs and not os.path.exists(dirs):
                os.makedirs(dirs)
            data = {}
            for key, val in self.output.items():
                if key.startswith('sol_'):
                    data[key] = [s.to_host().val for s in val]
                elif key.startswith('res_'):
                    data[key] = [{'info': r['info'], 'res_norm': r['cb'].res_norm} for r in val]
                else:
                    data[key] = val
            with open(filename, 'wb') as fop:
                pickle.dump(data, fop, protocol=3)

    def __repr__(self):
        return 'Class : Problem\n    name : %s\n    physics = %s\n    dim = %d (dimension)\n' % (
            getattr(self, 'name', ''), self.physics, self.dim)


def import_file(file_name):
    """Import an input file as a module (problem.py:143-160)."""
    module_path = os.path.dirname(os.path.abspath(file_name))
    remove = module_path not in sys.path
    if remove:
        sys.path.insert(0, module_path)
    try:
        module_name = os.path.splitext(os.path.basename(file_name))[0]
        conf = __import__(module_name)
    finally:
        if remove:
            sys.path.pop(0)
    return conf
