"""Copies the reference's IMC determinism fixtures (data only) into
tests/golden/imc: the keypoint positions/angles/sizes and RootSIFT
descriptors recorded on an RTX 3050 (determinism_reference/imc_*.npz,
manifest.json: n_octaves 4, S 3, sigma 1.6, ct 0.04, edge 10, fp16 pyramid,
DSP-5 descriptors) and the JPEG inputs they were computed from.  Descriptors
are stored as float16 (relative step 2^-11, far below the comparison's
cosine tolerance)."""
import os
import shutil
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "imc")
ITEMS = [
    ("imc_british_museum", "test_images/imc_phototourism/british_museum/set_100/images/00350405_2611802704.jpg"),
    ("imc_florence_cathedral_side",
     "test_images/imc_phototourism/florence_cathedral_side/set_100/images/05246669_3774806731.jpg"),
    ("imc_st_pauls_cathedral", "test_images/imc_phototourism/st_pauls_cathedral/21260922_4094863925.jpg"),
]

if __name__ == "__main__":
    os.makedirs(OUT, exist_ok=True)
    for name, img in ITEMS:
        z = np.load(os.path.join(REF, "determinism_reference", name + ".npz"))
        np.savez_compressed(os.path.join(OUT, name + ".npz"), points=z["points"].astype(np.float32),
                            angles=z["angles"].astype(np.float32), sizes=z["sizes"].astype(np.float32),
                            descriptors=z["descriptors"].astype(np.float16), image_shape=z["image_shape"])
        shutil.copyfile(os.path.join(REF, img), os.path.join(OUT, name + ".jpg"))
        print(name, len(z["points"]))
