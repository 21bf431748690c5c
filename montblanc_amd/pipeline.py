"""Kernel-node interface and ordered pipeline.

Same lifecycle and error behaviour as montblanc/node.py:21-139 and
montblanc/pipeline.py:29-177: ``initialise`` / ``pre_execution`` /
``execute`` / ``post_execution`` / ``shutdown(solver, stream=None)`` per node;
the pipeline logs and swallows exceptions. ``Pipeline.execute`` returns
False whatever happens, as the reference's does (pipeline.py:125-137); the
exception of a failed execution is kept in ``Pipeline.last_error`` so
callers can inspect it.
"""

import logging

log = logging.getLogger('montblanc')


class PipeLineError(Exception):
    pass


class Node(object):
    """Abstract pipeline node (node.py:21-126)."""

    def __init__(self):
        pass

    def description(self):
        return self.__class__.__name__

    def _missing(self, fn):
        return ' '.join(['method', fn, 'not implemented in class', type(self).__name__,
                         'derived from abstract class', Node.__name__])

    def initialise(self, solver, stream=None):
        raise NotImplementedError(self._missing('initialise'))

    def shutdown(self, solver, stream=None):
        raise NotImplementedError(self._missing('shutdown'))

    def execute(self, solver, stream=None):
        raise NotImplementedError(self._missing('execute'))

    def pre_execution(self, solver, stream=None):
        raise NotImplementedError(self._missing('pre_execution'))

    def post_execution(self, solver, stream=None):
        raise NotImplementedError(self._missing('post_execution'))


class NullNode(Node):
    """A node that does nothing (node.py:128-139)."""

    def initialise(self, solver, stream=None):
        pass

    def shutdown(self, solver, stream=None):
        pass

    def pre_execution(self, solver, stream=None):
        pass

    def execute(self, solver, stream=None):
        pass

    def post_execution(self, solver, stream=None):
        pass


class Pipeline(object):
    """Ordered list of nodes (pipeline.py:29-177)."""

    def __init__(self, node_list=None):
        if node_list is None:
            node_list = [NullNode()]
        if type(node_list) is not list:
            raise TypeError('node_list argument is not a list')
        self.pipeline = node_list
        self.initialised = False
        self.nr_of_executions = 0
        self.last_error = None

    def is_initialised(self):
        return self.initialised

    def initialise(self, solver, stream=None):
        if self.is_initialised():
            return True
        log.debug('Initialising pipeline')
        try:
            for node in self.pipeline:
                log.debug("Initialising node '%s'", node.description())
                node.initialise(solver, stream)
        except Exception as e:
            log.error('Exception during RIME pipeline initialisation', exc_info=True)
            self.last_error = e
            self.initialised = False
            return False
        self.initialised = True
        self.nr_of_executions = 0
        return True

    def execute(self, solver, stream=None):
        log.debug('Executing pipeline')
        if not self.is_initialised():
            log.error('Pipeline was not initialised!')
            if not self.initialise(solver, stream):
                return False
        self.last_error = None
        try:
            for node in self.pipeline:
                log.debug("Executing node '%s'", node.description())
                node.pre_execution(solver, stream)
                node.execute(solver, stream)
                node.post_execution(solver, stream)
            self.nr_of_executions += 1
        except Exception as e:
            log.error('Exception during RIME pipeline execution', exc_info=True)
            self.last_error = e
            return False
        log.debug('Execution of pipeline complete')
        return False

    def shutdown(self, solver, stream=None):
        log.debug('Shutting down pipeline')
        success = True
        for node in self.pipeline:
            try:
                node.shutdown(solver, stream)
            except Exception as e:
                log.error('Exception during RIME pipeline shutdown', exc_info=True)
                self.last_error = e
                success = False
        self.initialised = False
        return success

    def __str__(self):
        return ' '.join(node.description() for node in self.pipeline)
