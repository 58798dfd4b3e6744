"""Copy / extract the reference's data fixtures used by the tests (run in the build container, where the reference is
mounted at /root/reference).  Only data files: the PtNi cluster-expansion model file and the reference energies the
reference's fit script recorded for it.

    python tests/golden/make_fixtures.py
"""
import json
import os
import shutil

REF = os.environ.get("NCE_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    shutil.copyfile(f"{REF}/models/ce/ptni/icet_c8_6_5.ce", f"{HERE}/icet_c8_6_5.ce")
    fits = json.load(open(f"{REF}/models/ce/ptni/fits.json"))
    json.dump({"icet_c8_6_5": fits["icet_c8_6_5"]}, open(f"{HERE}/ptni_fits_c8_6_5.json", "w"), indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
