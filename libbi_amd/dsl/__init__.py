"""Front end for LibBi's model language: `parser` reads `.bi` files,
`codegen` checks them and writes the device functors that the library build
compiles in (libbi_amd/build.py runs it over libbi_amd/bi/*.bi)."""
from .codegen import GenError, analyse, generate, load_models  # noqa: F401
from .parser import BiSyntaxError, parse, parse_file  # noqa: F401
