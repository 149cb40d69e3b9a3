"""python -m residuum_hip {heateq,mesh} config.yaml"""
import sys

from .cli import main

sys.exit(main())
