"""Write config.mk for the host driver: the problem setup, EOS and flux policies, precision.

Same options as the reference's configure.py (configure.py:8-19: --setup, --eos, --flux,
-openmp) plus --precision.  The EOS and flux names select template instantiations of the HIP
stage kernels; -openmp is accepted for compatibility and ignored (the step runs on the GPU).
"""
import argparse
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SETUPS = sorted(f[6:-4] for f in os.listdir(os.path.join(HERE, "src", "setup"))
                if f.startswith("setup_") and f.endswith(".cpp"))


def main(argv=None):
    ap = argparse.ArgumentParser(description="Configure build options for the gaukuk_amd host driver")
    ap.add_argument("--setup", default="kh", choices=SETUPS)
    ap.add_argument("--eos", default="adiabatic", choices=["adiabatic"])
    ap.add_argument("--flux", default="hllc", choices=["hllc"])
    ap.add_argument("--precision", default="64", choices=["64", "32"])
    ap.add_argument("-openmp", action="store_true", help="accepted for compatibility; ignored")
    a = ap.parse_args(argv)
    text = (f"# configurable options\nSETUP = {a.setup}\nEOS = {a.eos}\nFLUX = {a.flux}\n"
            f"USE_SINGLE_PRECISION = {1 if a.precision == '32' else 0}\n")
    with open(os.path.join(HERE, "config.mk"), "w") as f:
        f.write(text)
    print(text, end="")


if __name__ == "__main__":
    main()
