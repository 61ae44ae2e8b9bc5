"""Checks model files and prints the code the library build would compile:
    python -m libbi_amd.dsl model.bi [more.bi ...]"""
import sys

from . import codegen
from .parser import BiSyntaxError


def main(argv):
    if not argv:
        print(__doc__, file=sys.stderr)
        return 2
    try:
        models = codegen.load_models(argv)
    except (BiSyntaxError, codegen.GenError) as e:
        print(f"error: {e}", file=sys.stderr)
        return 1
    sys.stdout.write(codegen.generate(models))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
