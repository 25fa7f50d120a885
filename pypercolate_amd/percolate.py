"""Drop-in for ``percolate.percolate`` (andsor/pypercolate 0.4.6).

Same functions, signatures, return values and errors as
percolate/percolate.py, with the sweep, the reduction over runs and the
canonical convolution running on the GPU (libpercolate_hip).  Differences
that follow from running on networkx 3 / numpy 2 and from exact integer
arithmetic are listed where they occur:

* graphs use networkx 3 (``graph.nodes[n]`` for ``graph.node[n]``), and edge
  order is ``perc_graph.edges()`` of the installed networkx;
* moments are exact integers converted to float64 at the end; the reference
  accumulates them in float64 (percolate.py:259,334-350), which agrees
  exactly while values stay below 2**53 (every lattice up to L = 128) and
  overflows int64 for clusters above 55108 nodes.
"""

import copy

import networkx as nx
import numpy as np
import scipy.stats

from . import _graph

alpha_1sigma = 2 * scipy.stats.norm.cdf(-1.0)
"""
The alpha for the 1 sigma confidence level
"""


def percolation_graph(graph, spanning_cluster=True):
    """Prepare the (internal) percolation graph from a given graph
    (percolate/percolate.py:31-100): strips the auxiliary nodes and returns
    the keyword dict the ``percolate.hpc`` functions take."""
    ret = dict()
    ret['graph'] = graph
    ret['spanning_cluster'] = bool(spanning_cluster)
    if spanning_cluster:
        spanning_auxiliary_node_attributes = nx.get_node_attributes(graph, 'span')
        ret['auxiliary_node_attributes'] = spanning_auxiliary_node_attributes
        if not list(spanning_auxiliary_node_attributes.keys()):
            raise ValueError(
                'Spanning cluster is to be detected, but no auxiliary nodes '
                'given.'
            )
        spanning_sides = list(set(spanning_auxiliary_node_attributes.values()))
        if len(spanning_sides) != 2:
            raise ValueError(
                'Spanning cluster is to be detected, but auxiliary nodes '
                'of less or more than 2 types (sides) given.'
            )
        ret['spanning_sides'] = spanning_sides
        ret['auxiliary_edge_attributes'] = nx.get_edge_attributes(graph, 'span')
        perc_graph = graph.subgraph(
            [node for node in graph.nodes() if 'span' not in graph.nodes[node]]
        )
    else:
        perc_graph = graph
    ret['perc_graph'] = perc_graph
    ret['num_nodes'] = nx.number_of_nodes(perc_graph)
    ret['num_edges'] = nx.number_of_edges(perc_graph)
    return ret


def spanning_1d_chain(length):
    """Linear chain with auxiliary nodes for spanning cluster detection
    (percolate/percolate.py:899-928)."""
    ret = nx.grid_graph(dim=[int(length + 2)])
    ret.nodes[0]['span'] = 0
    ret[0][1]['span'] = 0
    ret.nodes[length + 1]['span'] = 1
    ret[length][length + 1]['span'] = 1
    return ret


def spanning_2d_grid(length):
    """Square lattice with auxiliary nodes for spanning detection
    (percolate/percolate.py:931-965)."""
    ret = nx.grid_2d_graph(length + 2, length)
    for i in range(length):
        ret.nodes[(0, i)]['span'] = 0
        ret[(0, i)][(1, i)]['span'] = 0
        ret.nodes[(length + 1, i)]['span'] = 1
        ret[(length + 1, i)][(length, i)]['span'] = 1
    return ret


def _graph_arrays(graph, spanning_cluster):
    pg = percolation_graph(graph, spanning_cluster)
    return _graph.from_percolation_graph(**pg)


def _check_model_and_sides(graph, spanning_cluster, model):
    """The eager checks of percolate/percolate.py:210-227."""
    if model != 'bond':
        raise ValueError('Only bond percolation supported.')
    if spanning_cluster:
        aux = nx.get_node_attributes(graph, 'span')
        if not list(aux.keys()):
            raise ValueError(
                'Spanning cluster is to be detected, but no auxiliary nodes '
                'given.'
            )
        if len(set(aux.values())) != 2:
            raise ValueError(
                'Spanning cluster is to be detected, but auxiliary nodes '
                'of less or more than 2 types (sides) given.'
            )
