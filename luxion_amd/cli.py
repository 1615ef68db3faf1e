"""Command line renderer over the C ABI, with the reference's surface
(src/main.cpp:71-147, README.md "Usage"):

    luxion SCENEFILE [-e|--envmap ENVMAP] [-i|--iterations N] [-o|--output NAME] [-dmis] [-emis] [-lock]

Renders N samples per pixel (default 5000) of the .glb scene on one GPU and
writes NAME.png (default img/<scene name>.<UTC start time>.<N>samp.png) the
way the reference's saveImage does.  There is no window: -lock (start with
the camera locked) is accepted and has nothing to lock.
"""
import os
import sys
import time

USAGE = "Usage: %s SCENEFILE [-e|--envmap ENVMAP] [-i|--iterations N] [-o|--output NAME] [-dmis] [-emis] [-lock]"


class UsageError(Exception):
    pass


def parse_args(argv):
    """Returns dict(scene, envmap, iterations, output, dmis, emis, lock);
    raises UsageError with the reference's messages."""
    a = dict(scene=None, envmap=None, iterations=5000, output=None, dmis=False, emis=False, lock=False)
    i = 1
    while i < len(argv):
        x = argv[i]
        if x in ("-e", "--envmap"):
            if i + 1 >= len(argv):
                raise UsageError("Error: %s requires a path" % x)
            i += 1
            a["envmap"] = argv[i]
        elif x in ("-i", "--iterations"):
            if i + 1 >= len(argv):
                raise UsageError("Error: %s requires a number" % x)
            i += 1
            try:
                v = int(argv[i], 10)
            except ValueError:
                v = 0
            if v <= 0 or not argv[i].lstrip("+-").isdigit():
                raise UsageError("Error: invalid iteration count '%s'" % argv[i])
            a["iterations"] = v
        elif x in ("-o", "--output"):
            if i + 1 >= len(argv):
                raise UsageError("Error: %s requires a name" % x)
            i += 1
            a["output"] = argv[i]
        elif x == "-dmis":
            a["dmis"] = True
        elif x == "-emis":
            a["emis"] = True
        elif x == "-lock":
            a["lock"] = True
        elif a["scene"] is None:
            a["scene"] = x
        else:
            raise UsageError("Error: unexpected argument '%s'" % x)
        i += 1
    if a["scene"] is None:
        raise UsageError(USAGE % (argv[0] if argv else "luxion"))
    return a


def default_output(scene_name, start, spp):
    """img/<name>.<start>.<spp>samp (save_image.cpp:69-79); spp printed as the
    reference's float stream insertion prints it."""
    return "img/%s.%s.%ssamp" % (scene_name, start, "%g" % float(spp))


def main(argv=None):
    argv = list(sys.argv if argv is None else argv)
    try:
        a = parse_args(argv)
    except UsageError as e:
        print(e)
        return 1
    start = time.strftime("%Y-%m-%d_%H-%M-%Sz", time.gmtime())
    import luxion_amd as lx
    scene = a["scene"]
    if scene.endswith(".json"):
        print("Error: JSON scenes are deprecated; convert the scene to .glb")
        return 1
    if not scene.endswith(".glb"):
        print("Error: unsupported scene file '%s' (expected .glb)" % scene)
        return 1
    print("Loading %s as .glb file" % scene, flush=True)
    try:
        ctx = lx.Context(0)
        sc = lx.Scene.load_glb(ctx, scene, a["envmap"])
        r = lx.Render(sc, dmis=a["dmis"], emis=a["emis"])
    except lx.LuxionError as e:
        msg = str(e).split(": ", 1)[-1]
        print("Error: %s" % msg if not msg.startswith("-") else "Error: %s" % msg)
        return 1
    n = a["iterations"]
    info = sc.info()
    name = info["name"].decode() if isinstance(info["name"], bytes) else info["name"]
    t0 = time.perf_counter()
    done = 0
    chunk = max(1, min(64, n // 20 or 1))
    while done < n:
        k = min(chunk, n - done)
        r.iterate(done + 1, k)
        r.sync()
        done += k
        print("iteration %d / %d (%.1f spp/s)" % (done, n, done / (time.perf_counter() - t0)), flush=True)
    out = a["output"] or default_output(name, start, n)
    d = os.path.dirname(out)
    if d:
        os.makedirs(d, exist_ok=True)
    r.save_png(out + ".png")
    print("Saved %s.png." % out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
