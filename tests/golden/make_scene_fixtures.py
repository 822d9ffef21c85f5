"""Copies the reference's scene files (tests/configs/*.json, the input scenes of its
demo and test helper) into tests/golden/scenes/ as data fixtures, with the whitespace
outside strings removed. Numbers keep their exact text, so parsing the copy gives the
same values as parsing the original. Run from the repository root where the reference
is checked out at /root/reference."""
import glob
import os

SRC = "/root/reference/tests/configs"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scenes")


def minify(text):
    out, in_str, esc = [], False, False
    for ch in text:
        if in_str:
            out.append(ch)
            if esc:
                esc = False
            elif ch == "\\":
                esc = True
            elif ch == '"':
                in_str = False
        elif ch == '"':
            in_str = True
            out.append(ch)
        elif not ch.isspace():
            out.append(ch)
    return "".join(out)


def main():
    os.makedirs(DST, exist_ok=True)
    for path in sorted(glob.glob(os.path.join(SRC, "*.json"))):
        with open(path) as f:
            text = minify(f.read())
        with open(os.path.join(DST, os.path.basename(path)), "w") as f:
            f.write(text + "\n")


if __name__ == "__main__":
    main()
