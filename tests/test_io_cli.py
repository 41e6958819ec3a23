"""File formats and the command line (CPU parts)."""
import numpy as np
import pytest

from tomocupy_amd import io
from tomocupy_amd.cli import parser


@pytest.mark.parametrize("dt", [np.float32, np.float16])
def test_tiff_roundtrip(tmp_path, dt):
    rng = np.random.default_rng(0)
    img = rng.standard_normal((37, 53)).astype(dt)
    p = tmp_path / 'a.tiff'
    io.write_tiff(str(p), img)
    back = io.read_tiff(str(p))
    assert back.dtype == dt and np.array_equal(back, img)
    raw = p.read_bytes()
    assert raw[:4] == b'II*\x00'


def test_dataset_npz(tmp_path):
    rng = np.random.default_rng(1)
    data = rng.integers(0, 60000, (12, 5, 16), dtype=np.uint16)
    flat = rng.integers(0, 60000, (3, 5, 16), dtype=np.uint16)
    dark = rng.integers(0, 100, (2, 5, 16), dtype=np.uint16)
    p = tmp_path / 'd.npz'
    np.savez(p, data=data, data_white=flat, data_dark=dark, theta=np.linspace(0, 180, 12))
    ds = io.Dataset(str(p))
    assert (ds.nproj, ds.nz, ds.ni) == (12, 5, 16)
    assert np.allclose(ds.theta, np.linspace(0, np.pi, 12), atol=1e-6)
    d, k, f = ds.rows(1, 3)
    assert np.array_equal(d, data[:, 1:3]) and np.array_equal(k, dark[:, 1:3])
    assert np.array_equal(f, flat[:, 1:3])
    # no theta: linspace(0, pi, nproj) with the endpoint (reader.py:317)
    p2 = tmp_path / 'e.npz'
    np.savez(p2, data=data, data_white=flat, data_dark=dark)
    assert np.allclose(io.Dataset(str(p2)).theta, np.linspace(0, np.pi, 12))


def test_cli_defaults_match_reference():
    a = parser().parse_args(['recon', '--file-name', 'x.npz'])
    assert a.reconstruction_type == 'try' and a.fbp_filter == 'parzen'
    assert a.center_search_width == 50.0 and a.center_search_step == 0.5
    assert a.nsino_per_chunk == 8 and a.rotation_axis == -1.0 and a.dtype == 'float32'
    assert a.minus_log == 'True' and a.max_write_threads == 8


def test_console_entry_is_tomocupy():
    """setup.py installs the reference's command name (reference setup.py:9)."""
    import ast
    import os
    src = open(os.path.join(os.path.dirname(__file__), '..', 'setup.py')).read()
    call = next(n for n in ast.walk(ast.parse(src))
                if isinstance(n, ast.Call) and getattr(n.func, 'id', '') == 'setup')
    kw = {k.arg: ast.literal_eval(k.value) for k in call.keywords if k.arg == 'entry_points'}
    scripts = dict(s.replace(' ', '').split('=') for s in kw['entry_points']['console_scripts'])
    assert scripts['tomocupy'] == 'tomocupy_amd.cli:main'
    mod, fn = scripts['tomocupy'].split(':')
    import importlib
    main = getattr(importlib.import_module(mod), fn)
    with pytest.raises(SystemExit) as e:
        main(['recon', '--help'])
    assert e.value.code == 0


def test_cli_rejects_fw_sigma_beyond_notch_limit(capsys):
    from tomocupy_amd.cli import main
    with pytest.raises(SystemExit) as e:
        main(['recon', '--file-name', 'x.npz', '--remove-stripe-method', 'fw', '--fw-sigma', '12'])
    assert e.value.code == 2 and '--fw-sigma' in capsys.readouterr().err
