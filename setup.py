"""Installs the package with the reference's command name.

``tomocupy recon --reconstruction-algorithm fourierrec ...`` then runs this
package's CLI (the reference's setup.py:9 installs
``tomocupy = tomocupy.__main__:main``).  Build the HIP library in-tree first
(``python -m tomocupy_amd.build``); it ships as package data.
"""
from setuptools import setup

setup(
    name='tomocupy_amd',
    version='0.1.0',
    description="MI355X (gfx950) HIP implementation of tomocupy's fourierrec reconstruction path",
    packages=['tomocupy_amd'],
    package_data={'tomocupy_amd': ['libtomohip.so', 'csrc/*.hip', 'csrc/*.h']},
    entry_points={'console_scripts': ['tomocupy = tomocupy_amd.cli:main',
                                      'tomocupy_amd = tomocupy_amd.cli:main']},
    python_requires='>=3.9',
    install_requires=['numpy', 'torch'],
    zip_safe=False,
)
