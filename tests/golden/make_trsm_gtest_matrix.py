"""Write trsm_gtest_matrix.json: the reference's trsm test matrix
(clients/gtest/trsm_gtest.yaml with the clients/include/rocblas_common.yaml
and known_bugs.yaml it includes) as data -- the Defaults, the list-expansion
rules and every Tests entry with its anchors resolved, unexpanded.
rocblas_amd.clients expands it into cases exactly as it expands the YAML.

Run in a tree that has the reference at REF (default /root/reference):
    python tests/golden/make_trsm_gtest_matrix.py [REF]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from rocblas_amd.clients import load_yaml_doc  # noqa: E402


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    doc = load_yaml_doc(os.path.join(ref, "clients", "gtest", "trsm_gtest.yaml"),
                        [os.path.join(ref, "clients", "include")])
    out = {
        "source": "clients/gtest/trsm_gtest.yaml (+ clients/include/rocblas_common.yaml)",
        "Defaults": doc.get("Defaults") or {},
        "Dictionary lists to expand": doc.get("Dictionary lists to expand") or [],
        "Lists to not expand": doc.get("Lists to not expand") or [],
        "Tests": doc["Tests"],
    }
    path = os.path.join(HERE, "trsm_gtest_matrix.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"), sort_keys=True)
        f.write("\n")
    print(f"wrote {path}: {len(out['Tests'])} test entries, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
