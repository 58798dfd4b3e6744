"""Copy / extract the reference's data fixtures used by the tests (run in the build container, where the reference is
mounted at /root/reference).  Only data files and values the reference's own results hold:

  icet_c8_6_5.ce               PtNi cluster-expansion model (models/ce/ptni/icet_c8_6_5.ce)
  ptni_fits_c8_6_5.json        its reference-ordering energies (models/ce/ptni/fits.json, 8/6/5 block)
  nipt_v3_nomax_sel.pkl        the shipped NeuralCE model (models/nce/nipt_v3_nomax_sel/model.pkl), copied as bytes
  nce_features_subset.npz      layer-0 features, shell tables and model predictions of the 121 training cells
                               (results/cv/nipt_v3_nomax_sel/features.npz, fit_preds.npz): singlet column -> spins,
                               raw features H, one shell table (identical for all cells), predictions
  exact_N48.json               exact thermodynamics of the model on 48 sites (results/ptni_exact/exact_N48.json)
  init_L10_N48.cif             its L1_0 ground state (results/ptni_exact/ladder/init_L10_N48.cif)

    python tests/golden/make_fixtures.py
"""
import json
import os
import shutil

import numpy as np

REF = os.environ.get("NCE_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    shutil.copyfile(f"{REF}/models/ce/ptni/icet_c8_6_5.ce", f"{HERE}/icet_c8_6_5.ce")
    fits = json.load(open(f"{REF}/models/ce/ptni/fits.json"))
    json.dump({"icet_c8_6_5": fits["icet_c8_6_5"]}, open(f"{HERE}/ptni_fits_c8_6_5.json", "w"), indent=1)
    shutil.copyfile(f"{REF}/models/nce/nipt_v3_nomax_sel/model.pkl", f"{HERE}/nipt_v3_nomax_sel.pkl")
    F = np.load(f"{REF}/results/cv/nipt_v3_nomax_sel/features.npz", allow_pickle=False)
    P = np.load(f"{REF}/results/cv/nipt_v3_nomax_sel/fit_preds.npz", allow_pickle=False)
    assert all(np.array_equal(F["SI"][0], F["SI"][i]) for i in range(len(F["SI"])))
    assert list(P["ids"]) == list(F["ids"])
    np.savez_compressed(f"{HERE}/nce_features_subset.npz", spins=np.rint(F["H"][:, :, 1]).astype(np.int8),
                        H=F["H"], SI=F["SI"][0], SM=F["SM"][0], pred=P["pred"], y=P["y"], ids=F["ids"])
    e = json.load(open(f"{REF}/results/ptni_exact/exact_N48.json"))
    json.dump({k: e[k] for k in ("T", "E", "Cv", "N", "nPt", "E_min", "E_max", "E_L10", "w_total")},
              open(f"{HERE}/exact_N48.json", "w"))
    shutil.copyfile(f"{REF}/results/ptni_exact/ladder/init_L10_N48.cif", f"{HERE}/init_L10_N48.cif")
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
