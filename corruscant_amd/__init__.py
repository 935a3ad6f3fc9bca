"""MI355X pair-counting path of corruscant: same Python API, GPU back end.

The two native libraries are found the way the reference finds its own
(corruscant/__init__.py:9-20): the first file in ``corruscant_amd/bin`` whose
name matches ``libkdtree`` / ``libcoords`` is opened with ctypes.  There is no
CPU fallback: without the gfx950 libraries the import fails.
"""
from ctypes import CDLL
import os
import re

PROJECT_PATH = os.path.dirname(os.path.abspath(__file__))


def loadlib(name):
    pattern = re.compile(name)
    binpath = os.path.join(PROJECT_PATH, "bin")
    names = sorted(os.listdir(binpath)) if os.path.isdir(binpath) else []
    for filename in names:
        if re.match(pattern, filename):
            return CDLL(os.path.join(binpath, filename))
    raise OSError("Shared library {} not found in {}: build it with "
                  "`make -C corruscant_amd/csrc` (or __graft_entry__.build())".format(name, binpath))


kdlib = loadlib("libkdtree")
coordlib = loadlib("libcoords")
