"""The configuration front end on the CPU: the reference's XML files, Python
initial data (init_global and per-cell forms), the reference sampler's draw
order with several parameters, error behaviour, and the NetCDF writer's file
contents read back by an independent reader (scipy.io.netcdf_file)."""
import os

import numpy as np
import pytest
import scipy.io

from alsvinn_amd import config, initial_data as ic, netcdf, sampling
from alsvinn_amd.sampling import MT19937_64, canonical

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs")


def test_read_kh_uq_config():
    c = config.read_config(os.path.join(HERE, "kh_uq.xml"))
    assert c.dims == (32, 32, 1) and c.lower == (0, 0, 0) and c.upper == (1, 1, 0)
    assert (c.equation, c.flux, c.reconstruction, c.integrator, c.boundary) == \
        ("euler2", "hll3", "weno2", "rungekutta3", "periodic")
    assert c.cfl == 0.475 and c.end_time == 0.06 and c.gamma == 1.4
    assert c.samples == 3 and c.has_uq and [p.name for p in c.uq_parameters] == ["a"]
    assert [s.name for s in c.stats] == ["meanvar", "structure_cube", "onepointmoment", "bv"]
    assert c.stats[1].params == {"numberOfH": 5.0, "p": 2.0}
    assert c.writer.number_of_saves == 2 and c.writer.basename == "kh_uq"


def _copy_config(tmp_path, text):
    import shutil
    shutil.copy(os.path.join(HERE, "kh_uq.py"), str(tmp_path / "kh_uq.py"))
    path = str(tmp_path / "kh_uq.xml")
    open(path, "w").write(text)
    return path


def test_auto_rules_follow_the_reference(tmp_path):
    """SimulatorSetup.cpp readCFLNumber / readIntegrator: none -> 0.9 and
    forward Euler; otherwise 0.475 and SSP-RK2."""
    c = config.read_config(os.path.join(HERE, "sod_snippet.xml"))
    assert c.cfl == 0.9 and c.integrator == "forwardeuler" and not c.has_uq
    src = open(os.path.join(HERE, "kh_uq.xml")).read().replace(
        "<integrator>rungekutta3</integrator>", "<integrator>auto</integrator>")
    assert config.read_config(_copy_config(tmp_path, src)).integrator == "rungekutta2"


@pytest.mark.parametrize("change, message", [
    ("<boundary>periodic</boundary>", "Unsupported configuration option bogus"),
    ("<flux>hll3</flux>", "flux rusanov is not available for euler2"),
    ("<flux>hll3</flux>", "flux tecno1 is not available on the hip platform"),
    ("<type>netcdf</type>\n    <basename>kh_uq</basename>\n    <numberOfSaves>2",
     "writer type hdf5 is not available"),
])
def test_rejected_options(change, message, tmp_path):
    src = open(os.path.join(HERE, "kh_uq.xml")).read()
    repl = {"<boundary>periodic</boundary>": "<boundary>periodic</boundary><bogus>1</bogus>",
            "<flux>hll3</flux>": "<flux>%s</flux>" % message.split()[1],
            "<type>netcdf</type>\n    <basename>kh_uq</basename>\n    <numberOfSaves>2":
                "<type>hdf5</type>\n    <basename>kh_uq</basename>\n    <numberOfSaves>2"}[change]
    assert change in src
    path = _copy_config(tmp_path, src.replace(change, repl, 1))
    with pytest.raises(config.ConfigError, match=message):
        config.read_config(path)


def test_wenof2_is_euler_only(tmp_path):
    """ReconstructionFactory.cpp:102-117: wenof2 exists for euler1/2/3 only."""
    src = open(os.path.join(HERE, "kh_uq.xml")).read().replace(
        "<reconstruction>weno2</reconstruction>", "<reconstruction>wenof2</reconstruction>")
    assert config.read_config(_copy_config(tmp_path, src)).reconstruction == "wenof2"
    src = open(os.path.join(HERE, "sod_snippet.xml")).read()
    src = src.replace("<equation>euler1</equation>", "<equation>burgers</equation>").replace(
        "<flux>hll</flux>", "<flux>godunov</flux>").replace(
        "<reconstruction>none</reconstruction>", "<reconstruction>wenof2</reconstruction>")
    with pytest.raises(config.ConfigError, match="WENO2F for Burgers"):
        config.read_config(_copy_config(tmp_path, src))


def test_onepoint_and_bv_need_integer_p(tmp_path):
    """OnePointMoment and BoundedVariation read p with getInteger."""
    src = open(os.path.join(HERE, "kh_uq.xml")).read()
    c = config.read_config(os.path.join(HERE, "kh_uq.xml"))
    assert [(s.name, s.params.get("p")) for s in c.stats][2:] == [("onepointmoment", 2.0),
                                                                  ("bv", 1.0)]
    bad = src.replace("<name>bv</name>\n      <numberOfSaves>1</numberOfSaves>\n      <p>1</p>",
                      "<name>bv</name>\n      <numberOfSaves>1</numberOfSaves>\n      <p>1.5</p>")
    assert bad != src
    with pytest.raises(config.ConfigError, match="integer <p>"):
        config.read_config(_copy_config(tmp_path, bad))


def test_init_global_matches_library_kelvin_helmholtz():
    """The configuration's init_global script (parameters as globals, arrays
    (nx, ny, nz) x fastest) gives the same conserved volume, bit for bit, as
    the library's Kelvin-Helmholtz initial data for the same draws."""
    c = config.read_config(os.path.join(HERE, "kh_uq.xml"))
    P = config.sample_parameters(c)
    X = sampling.stl_mersenne_uniform(3, 40)
    for i in range(3):
        assert np.array_equal(P[i]["a"], X[i])
        v = config.conserved_volume(c, 2, P[i])
        ref = ic.kelvin_helmholtz_2d(32, X[i])
        assert np.array_equal(v.view(np.uint64), ref.view(np.uint64))


def test_snippet_matches_library_sod():
    c = config.read_config(os.path.join(HERE, "sod_snippet.xml"))
    v = config.conserved_volume(c, 1)
    assert np.array_equal(v.view(np.uint64), ic.sod(128).view(np.uint64))


def _nth(indices):
    eng = MT19937_64()
    raw = eng.raw(max(indices) + 1)
    return canonical(raw[list(indices)])


def test_sampler_draw_order_with_two_parameters():
    """STLMersenne.cpp:36-85 traced by hand for parameters a (length 2) and
    b (length 3), sorted by name: the shared draw counter skips to
    sample*dimension and to the component, so sample 0 takes draws a=[0,1],
    b=[2,4,5] and sample 1 a=[6,7], b=[8,10,11] of the process-wide
    std::mt19937_64."""
    c = config.Config(name="x", platform="hip", lower=(0, 0, 0), upper=(1, 1, 0), dims=(4, 4, 1),
                      boundary="periodic", flux="hll", end_time=1.0, equation="euler2",
                      equation_parameters={}, reconstruction="none", cfl=0.9,
                      integrator="forwardeuler", python_file="", python_program="",
                      initial_parameters={}, samples=2, has_uq=True,
                      uq_parameters=[config.UQParameter("b", 3, "uniform"),
                                     config.UQParameter("a", 2, "uniform")])
    P = config.sample_parameters(c)
    assert np.array_equal(P[0]["a"], _nth([0, 1]))
    assert np.array_equal(P[0]["b"], _nth([2, 4, 5]))
    assert np.array_equal(P[1]["a"], _nth([6, 7]))
    assert np.array_equal(P[1]["b"], _nth([8, 10, 11]))


def test_netcdf_writer_reads_back_with_reference_layout(tmp_path):
    """dims member, time, x, y; variables rho, E, mx, my (+ coordinates);
    a slab holds the interior cells in memory order (x fastest) as the
    reference's NetCDFWriter::write copies them."""
    rng = np.random.default_rng(5)
    nx, ny, G = 6, 4, 2
    dims = (nx, ny, 1)
    path = str(tmp_path / "out")
    w = netcdf.EnsembleWriter(path, dims, netcdf.midpoints((0, 0, 0), (1, 2, 0), dims), 3,
                              netcdf.save_fractions(2, True), netcdf.variable_names(4, 2))
    vols = rng.random((3, 3, 4, 1, ny + 2 * G, nx + 2 * G))
    for m in range(3):
        for t in range(3):
            w.write(m, t, vols[m, t], (G, G, 0))
    f = scipy.io.netcdf_file(path, "r", mmap=False)
    try:
        assert f.dimensions["time"] == 3 and f.dimensions["x"] == nx and f.dimensions["y"] == ny
        assert list(f.variables["member"][:]) == [0, 1, 2]
        assert np.allclose(f.variables["time"][:], [0.0, 0.5, 1.0])
        assert np.allclose(f.variables["x"][:], (np.arange(nx) + 0.5) / nx)
        assert np.allclose(f.variables["y"][:], 2 * (np.arange(ny) + 0.5) / ny)
        for name, var in (("rho", 0), ("E", 3), ("mx", 1), ("my", 2)):
            data = f.variables[name][:]
            assert data.shape == (3, 3, nx, ny)
            for m in range(3):
                for t in range(3):
                    inner = vols[m, t, var, 0, G:-G, G:-G]
                    assert np.array_equal(data[m, t], inner.ravel().reshape(nx, ny))
    finally:
        f.close()
