"""Copies the reference's committed known-answer vector files for this path
into tests/golden/ (data only: inputs and expected outputs).

Small files are copied whole. The two large curve files are cut to a prefix of
each record kind (every header, edge and rejection line is kept; the random
`fq_ops` / `g1_ops` records are kept up to a count), which keeps the fixtures
small while every record kind stays covered.

Run in a checkout that has the reference at /root/reference:
    python tests/golden/extract_fixtures.py
"""
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

WHOLE = {
    "crates/curve/tests/vectors/msm_kats.txt": "msm_kats.txt",
    "crates/field/tests/vectors/fr_kats.txt": "fr_kats.txt",
    "crates/poly/tests/vectors/poly_kats.txt": "poly_kats.txt",
    "crates/transcript/tests/vectors/poseidon2_perm.txt": "poseidon2_perm.txt",
    "crates/transcript/tests/vectors/transcript_cases.txt": "transcript_cases.txt",
    "crates/pcs/tests/vectors/g1_absorb_kats.txt": "g1_absorb_kats.txt",
    "crates/pcs/tests/vectors/mercury_proof.txt": "mercury_proof.txt",
    "crates/pcs/tests/vectors/mercury_batch.txt": "mercury_batch.txt",
}
CUT = {
    "crates/curve/tests/vectors/fq_kats.txt": ("fq_kats_subset.txt", {"fq_ops": 300, "fq2_ops": 0}),
    "crates/curve/tests/vectors/g1_kats.txt": ("g1_kats_subset.txt", {"g1_ops": 120}),
}


def main():
    for src, dst in WHOLE.items():
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))
    for src, (dst, limits) in CUT.items():
        seen = {}
        out = []
        with open(os.path.join(REF, src)) as f:
            for line in f:
                kind = line.split(" ", 1)[0].strip()
                if kind in limits:
                    seen[kind] = seen.get(kind, 0) + 1
                    if seen[kind] > limits[kind]:
                        continue
                out.append(line)
        with open(os.path.join(HERE, dst), "w") as f:
            f.write(f"# subset of the reference's {src}; see extract_fixtures.py\n")
            f.writelines(out)


if __name__ == "__main__":
    main()
