"""``python -m create_rate_subs_amd -l hip -i mech.dat -t therm.dat``."""
import sys

from .create_rate_subs import main

sys.exit(main())
