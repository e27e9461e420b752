"""Writes headless_keys.json: the key order of the reference's two headless JSON lines, the
"headlessBenchmark" line of runBenchmark and the "backendComparison" line of
runBackendComparison (headless_server.cpp:605-943), read from the reference's source text."""
import json
import os
import re

SRC = "/root/reference/source/headless_server.cpp"
HERE = os.path.dirname(os.path.abspath(__file__))


def keys_of(text, start, end):
    body = text[text.index(start):text.index(end)]
    return re.findall(r'out << "\\"(\w+)\\":', body)


def main():
    text = open(SRC).read()
    out = {"headlessBenchmark": keys_of(text, "int runBenchmark(", "int runBackendComparison("),
           "backendComparison": keys_of(text, "int runBackendComparison(", "int main(")}
    with open(os.path.join(HERE, "headless_keys.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print({k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
