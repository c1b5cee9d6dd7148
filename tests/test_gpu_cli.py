"""The drop-in CURVELET binary end to end on the GPU: same invocation as the
reference's (run from a directory whose ../test_files holds the input), same
printed counts, and chain/info files byte-identical to the reference CLI's.
"""
import gzip
import json
import re
import subprocess
from pathlib import Path

import pytest

from tests._inputs import GOLDEN, shipped_text

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "curvelet_construction_on_gpu_amd" / "bin" / "CURVELET"

CASES = [("digit1_G4", "TO_edges_digit1", 4, []), ("digit1_G7", "TO_edges_digit1", 7, ["--output", "chain_g7.txt"]),
         ("abc_G4", "TO_edges_ABC_0006_thresh1", 4, ["--double"]), ("toedges_G7", "TO_edges", 7, [])]


@pytest.mark.parametrize("case,inp,G,extra", CASES)
def test_cli_outputs_identical(tmp_path, case, inp, G, extra):
    assert CLI.exists(), "CURVELET is not built"
    (tmp_path / "test_files").mkdir()
    (tmp_path / "run").mkdir()
    (tmp_path / "test_files" / f"{inp}.txt").write_bytes(shipped_text(inp))
    args = [str(CLI), "--edge-file", f"{inp}.txt", "--group-max-sz", str(G), "--nthreads", "4"] + extra
    r = subprocess.run(args, cwd=tmp_path / "run", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    meta = json.loads((GOLDEN / "ref_cli" / f"{case}.json").read_text())
    m = re.search(r"CSR neighbor graph: (\d+) anchors, (\d+) total pairs, max number of neighbors per anchor: (\d+)",
                  r.stdout)
    assert m and (int(m.group(1)), int(m.group(2)), int(m.group(3))) == (meta["edges"], meta["pairs"],
                                                                          meta["max_deg"])
    assert f"Pairwise curve bundles formed: {meta['valid_pairs']} valid pairs" in r.stdout
    assert f"(out_h, out_w) = ({meta['out_h']}, {meta['out_w']})" in r.stdout
    chain_name = "chain_cpu_double.txt" if "--double" in extra else (
        extra[extra.index("--output") + 1] if "--output" in extra else "chain_cpu.txt")
    info_name = chain_name.replace("chain", "info", 1)
    out = tmp_path / "outputs"
    chain, info = (out / chain_name).read_bytes(), (out / info_name).read_bytes()
    import hashlib
    assert hashlib.sha256(chain).hexdigest() == meta["chain_sha256"]
    assert hashlib.sha256(info).hexdigest() == meta["info_sha256"]
    gz = GOLDEN / "ref_cli" / f"{case}_chain.txt.gz"
    if gz.exists():
        with gzip.open(gz) as f:
            assert chain == f.read()
