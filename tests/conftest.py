import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def kb_fixture():
    import kinship_batch_fixture as kb
    return kb.make_bed(), kb.N_SAMPLES, kb.N_SNPS, kb.selections()


@pytest.fixture(scope="session")
def ref_datasets():
    import oracle
    g = os.path.join(ROOT, "tests", "golden")
    out = {}
    for name, prefix in [("synth", "ref_synth/synth"), ("gemma", "ref_gemma_accuracy/gemma_accuracy"),
                         ("ebfmt", "ref_ebfmt/ebfmt")]:
        out[name] = oracle.read_bed(os.path.join(g, prefix))
    return out
