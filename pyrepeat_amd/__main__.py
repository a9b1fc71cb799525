import sys

from .repeat import main

sys.exit(main())
