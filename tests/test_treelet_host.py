"""The GPU treelet-restructuring code, compiled for the CPU, against the restatement.

The region of `bisemutum-engine_amd/csrc/lbvh.hip` between its host-wave markers (one
wavefront's treelet formation, split search and rebuild) is compiled with g++ together
with `tests/host_wave/treelet_host.cpp`, which runs it with 64 threads as the lanes of one
wavefront, meeting at every wave barrier and shuffle.  Starting from the restatement's
Morton/Karras tree, its trees after 1 and 2 rounds must equal the restatement's
(`oracle/ref32.cpp` treelet_process) bit for bit.  This pins the kernel's source logic on
the CPU; the GPU parity tests pin the compiled kernel.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "bisemutum-engine_amd", "csrc", "lbvh.hip")


@pytest.fixture(scope="module")
def host_binary(tmp_path_factory):
    text = open(SRC).read()
    a = text.index("// [host-wave begin]")
    b = text.index("// [host-wave end]")
    d = tmp_path_factory.mktemp("host_wave")
    body = d / "treelet_body.inc"
    body.write_text(text[a:b])
    exe = d / "treelet_host"
    subprocess.run(["g++", "-O1", "-std=c++20", "-pthread", "-ffp-contract=off", f'-DTREELET_BODY="{body}"',
                    "-o", str(exe), os.path.join(HERE, "host_wave", "treelet_host.cpp")], check=True)
    return exe


def run_host(exe, tmp_path, bvh0, rounds):
    n = len(bvh0["order"])
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        np.array([n], np.int32).tofile(f)
        bvh0["children"].astype(np.int32).tofile(f)
        bvh0["bounds"].astype(np.float32).tofile(f)
        bvh0["parents"].astype(np.int32).tofile(f)
    subprocess.run([str(exe), str(src), str(dst), str(rounds)], check=True, timeout=120)
    raw = np.fromfile(dst, np.int32)
    ni = n - 1
    return raw[:2 * ni].reshape(ni, 2), raw[2 * ni:].view(np.float32).reshape(ni, 12)


@pytest.mark.parametrize("n", [12, 40, 200, 1000])
def test_treelet_kernel_source_matches_restatement(oracle_lib, host_binary, tmp_path, n):
    # (64 threads meeting at every barrier: trees of up to ~100 treelets keep this quick)
    sc = scenes.scene_random_triangles(n, 7)
    karras = oracle_lib.OracleScene(sc)
    karras.set_treelet_rounds(0)
    bvh0 = karras.bvh()
    for rounds in (1, 2):
        ch, bo = run_host(host_binary, tmp_path, bvh0, rounds)
        ref = oracle_lib.OracleScene(sc)
        ref.set_treelet_rounds(rounds)
        want = ref.bvh()
        assert np.array_equal(ch, want["children"]), (n, rounds)
        assert np.array_equal(bo.view(np.uint32), want["bounds"].view(np.uint32)), (n, rounds)
        if rounds == 1:
            assert not np.array_equal(ch, bvh0["children"])   # something was restructured
