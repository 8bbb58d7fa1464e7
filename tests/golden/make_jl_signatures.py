"""Write jl_signatures.json: the keyword names of the reference functions the
Julia shim replaces, read from the reference's source text (no Julia needed).

    python tests/golden/make_jl_signatures.py /root/reference

Functions (path:line in DiffMoM.jl): assemble_Z_efie
(src/assembly/EFIE.jl:699-710), matrixfree_efie_operator (:975-984),
solve_scattering (src/Workflow.jl:67-88), solve_gmres
(src/solver/IterativeSolve.jl:341-392), assemble_multiple_excitations
(src/assembly/Excitation.jl:2738-2797)."""
import json
import os
import re
import sys

FUNCS = {"assemble_Z_efie": "src/assembly/EFIE.jl", "matrixfree_efie_operator": "src/assembly/EFIE.jl",
         "solve_scattering": "src/Workflow.jl", "solve_gmres": "src/solver/IterativeSolve.jl",
         "assemble_multiple_excitations": "src/assembly/Excitation.jl"}


def keywords(src, name):
    """Keyword names of the first method `function name(...; kw...)` in src."""
    m = re.search(r"^function " + re.escape(name) + r"\((.*?)\)\s*$", src, flags=re.S | re.M)
    if not m:
        raise SystemExit(f"{name} not found")
    sig = m.group(1)
    depth, cut = 0, None
    for i, ch in enumerate(sig):
        depth += ch in "([{"
        depth -= ch in ")]}"
        if ch == ";" and depth == 0:
            cut = i
            break
    if cut is None:
        return []
    kw = sig[cut + 1:]
    names, depth, tok = [], 0, ""
    for ch in kw + ",":
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            t = tok.strip()
            if t:
                names.append(re.match(r"(\w+)", t).group(1))
            tok = ""
        else:
            tok += ch
    return names


def main(ref):
    out = {}
    for name, path in FUNCS.items():
        out[name] = keywords(open(os.path.join(ref, path)).read(), name)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jl_signatures.json")
    json.dump(out, open(dst, "w"), indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
