"""Result rows as the reference's command-line program writes them (main.cpp:24-55): one CSV line
per outer iteration of the last run, appended to <csv><num_vertices>/<type>.csv, with the header
written when the file is new. Numbers are printed as C++ streams print them by default (six
significant digits, %g); booleans as 0/1."""
import os

HEADER = ("tag, numvertices, numedges, overlap, blocksizevar, undirected, algorithm, iteration, mdl, "
          "normalized_mdl_v1, normalized_mdl_v2, modularity, interblock_edges, block_size_variation, f1_score, "
          "nmi, true_mdl, true_mdl_v1, true_mdl_v2, runtime, mcmc_iterations")


def _num(x):
    """A double or float as std::ostream prints it with its default format."""
    x = float(x)
    if x != x:
        return "nan" if str(x)[0] != "-" else "-nan"
    if x in (float("inf"), float("-inf")):
        return "inf" if x > 0 else "-inf"
    return "%g" % x


def null_mdl_v1(E):
    """entropy::null_mdl_v1 (entropy.cpp:567-573)."""
    import math

    log_posterior_p = E * math.log(1.0 / E)
    x = 1.0 / E
    h = ((1 + x) * math.log(1 + x)) - (x * math.log(x))
    return (E * h) - log_posterior_p


def null_mdl_v2(N, E):
    """entropy::null_mdl_v2 (entropy.cpp:575-582)."""
    import math

    log_posterior_p = E * math.log(1.0 / E)
    x = (N ** 2) / E
    h = ((1 + x) * math.log(1 + x)) - (x * math.log(x))
    return (E * h) + (N * math.log(N)) - log_posterior_p


def write_results(csv, type_, num_vertices, num_edges, intermediates, f1_score, nmi, true_mdl, runtime,
                  tag="", overlap="", blocksizevar="", undirected=False, algorithm="async_gibbs", modularity=None):
    """Appends the run's rows (`intermediates` as ResidentGraph.intermediates() returns them) to
    f"{csv}{num_vertices}/{type_}.csv". `modularity` fills the final row's (iteration -1) column,
    which the reference computes only there. Returns the file path."""
    directory = f"{csv}{num_vertices}"
    path = f"{directory}/{type_}.csv"
    exists = os.path.exists(path)
    os.makedirs(directory, exist_ok=True)
    v1 = true_mdl / null_mdl_v1(num_edges)
    v2 = true_mdl / null_mdl_v2(num_vertices, num_edges)
    with open(path, "a") as f:
        if not exists:
            f.write(HEADER + "\n")
        for r in intermediates:
            q = r["modularity"]
            if r["iteration"] == -1 and modularity is not None:
                q = modularity
            fields = [str(tag), str(num_vertices), str(num_edges), str(overlap), str(blocksizevar),
                      "1" if undirected else "0", algorithm, _num(r["iteration"]), _num(r["mdl"]),
                      _num(r["normalized_mdl_v1"]), _num(r["normalized_mdl_v2"]), _num(q),
                      _num(r["interblock_edges"]), _num(r["block_size_variation"]), _num(f1_score), _num(nmi),
                      _num(true_mdl), _num(v1), _num(v2), _num(runtime), str(int(r["mcmc_iterations"]))]
            f.write(", ".join(fields) + "\n")
    return path
