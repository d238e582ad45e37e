"""INQ's text checkpoint values (utils::save_value / save_container,
src/utils/load_save.hpp): 25 significant digits, one value per line, vector3
as tab-separated components; real_time::results round trip."""
import os

import numpy as np

from inq_amd import checkpoint


def test_real_time_results_text_format(tmp_path):
    d = str(tmp_path / "observables")
    t = [0.0, 0.1, 0.2]
    e = [-1.5, -1.5000000001, -1.4999999999]
    cur = [(0.0, 0.0, 0.0), (1e-3, -2e-3, 0.5), (2e-3, -4e-3, 1.0)]
    checkpoint.save_real_time_results(d, 2, 0.2, t, e, current=cur)
    assert open(os.path.join(d, "total_steps")).read() == "2\n"
    assert open(os.path.join(d, "total_time")).read() == "0.2000000000000000111022302\n"
    lines = open(os.path.join(d, "time")).read().splitlines()
    assert lines == ["0", "0.1000000000000000055511151", "0.2000000000000000111022302"]
    cl = open(os.path.join(d, "current")).read().splitlines()
    assert cl[1].split("\t") == ["%.25g" % 1e-3, "%.25g" % -2e-3, "0.5"]
    # an empty container leaves no file (save_container, load_save.hpp:102-106)
    assert not os.path.exists(os.path.join(d, "dipole"))
    r = checkpoint.load_real_time_results(d)
    assert r["total_steps"] == 2 and r["time"] == t and r["total_energy"] == e
    assert np.array_equal(np.array(r["current"]), np.array(cur)) and r["dipole"] == []
