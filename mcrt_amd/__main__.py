"""``python -m mcrt_amd --headless <scene.yaml> --output <png> [--width W]
[--height H] [--spp N] [--denoise bilateral]`` (src/main.rs:56-80)."""
import sys

from .headless import main

sys.exit(main())
