"""Command-line entry of the headless benchmark (the reference's avbd_headless_server benchmark
mode, headless_server.cpp:946-985); see avbd/headless.py."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from avbd.headless import main  # noqa: E402

if __name__ == "__main__":
    sys.exit(main())
