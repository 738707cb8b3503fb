"""Entry point with the reference's command line: see chemgen_amd/generate_chemistry.py."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from chemgen_amd.generate_chemistry import main  # noqa: E402

if __name__ == "__main__":
    main()
