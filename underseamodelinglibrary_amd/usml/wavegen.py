"""usml::wavegen over libwq3d.so: wavefront_generator tasks and a batch
runner that propagates many of them at once on one GPU.

wavefront_generator mirrors the reference task (wavegen/wavefront_generator.h:
38-130, wavefront_generator.cc:27-127): it propagates one source's fan to
time_maximum, collects eigenrays (eigenray_collection, summed at the end)
and eigenverbs (eigenverb_collection), optionally writes the wavefront log,
and hands both collections to the source's wavefront listeners.  abort()
stops it before or during the run, as thread_task::abort does.

The reference runs each generator as a thread_task on a CPU thread_pool
(threads/thread_pool.cc:21-75).  Here wavefront_batch plays the pool's part
for the GPU: every generator's fan lives on its own HIP stream, the batch
queues a chunk of steps on every live fan before it waits for any of them,
so the fans' kernels overlap on the device, and then replays each fan's
listener calls in the reference's per-step order.  Each fan's results are
the same as a run on its own.  Abort is checked after every chunk (after
every step when the generator writes a wavefront log).

The source is any object with the parts of sensors/sensor_model.h the task
uses: position(), keyID(), compute_reverb(), description() and
notify_wavefront_listeners(source, eigenrays, eigenverbs).  wavefront_source
is a plain one.
"""
import itertools

import numpy as np

from .eigenrays import eigenray_collection


class ocean_shared:
    """ocean/ocean_shared.h: the process-wide ocean that new generators
    take at construction."""
    _current = None

    @classmethod
    def current(cls):
        return cls._current

    @classmethod
    def update(cls, ocean):
        cls._current = ocean


class eigenverb_collection:
    """eigenverbs/eigenverb_collection.h:55-140, the storage part: eigenverbs
    filed per interface (bottom 0, surface 1, volume layer l upper 2 + 2l,
    lower 3 + 2l) in arrival order.  The spatial query (find_eigenverbs)
    is not provided."""

    def __init__(self, num_volumes=0):
        self._collection = [[] for _ in range(2 + 2 * int(num_volumes))]

    def num_interfaces(self):
        return len(self._collection)

    def size(self, interface):
        return len(self._collection[interface])

    def eigenverbs(self, interface):
        return list(self._collection[interface])

    def add_eigenverb(self, verb, interface):
        self._collection[interface].append(verb)


class wavefront_source:
    """A source with the sensor_model members the generator reads."""

    def __init__(self, position, keyID=0, compute_reverb=False, description="source"):
        self._position = position
        self._key = int(keyID)
        self._reverb = bool(compute_reverb)
        self._description = description
        self.results = []   # (eigenrays, eigenverbs) per finished generator

    def position(self):
        return self._position

    def keyID(self):
        return self._key

    def compute_reverb(self):
        return self._reverb

    def description(self):
        return self._description

    def notify_wavefront_listeners(self, source, eigenrays, eigenverbs):
        self.results.append((eigenrays, eigenverbs))


def _steps_to(time_step, time_maximum):
    """Number of step() calls of `while (wave.time() < time_maximum)` with
    time accumulated by repeated addition, as wave_queue keeps it."""
    t, n = 0.0, 0
    while t < time_maximum:
        t += time_step
        n += 1
    return n


class wavefront_generator:
    """wavegen/wavefront_generator.h:38-130."""

    _ids = itertools.count(1)

    def __init__(self, source, target_positions, targetIDs, frequencies, de_fan, az_fan,
                 time_step, time_maximum, intensity_threshold, max_bottom, max_surface,
                 wavefront_file="", queue_factory=None, ocean=None, collection_lib=None,
                 verbose=False):
        """The reference's arguments, then: queue_factory builds the fan
        (default: the device wave_queue), ocean replaces ocean_shared's,
        collection_lib is the library object the eigenray collection sums
        with (default: libwq3d)."""
        self._ocean = ocean if ocean is not None else ocean_shared.current()
        if self._ocean is None:
            raise ValueError("no ocean: call ocean_shared.update() or pass ocean=")
        self._source = source
        self._source_position = source.position()
        self._target_positions = target_positions
        self._targetIDs = None if targetIDs is None else np.asarray(targetIDs, dtype=np.uint64)
        self._frequencies = frequencies
        self._de_fan = de_fan
        self._az_fan = az_fan
        self._time_step = float(time_step)
        self._time_maximum = float(time_maximum)
        self._intensity_threshold = float(intensity_threshold)
        self._max_bottom = int(max_bottom)
        self._max_surface = int(max_surface)
        self._wavefront_file = wavefront_file or ""
        self._factory = queue_factory
        self._collection_lib = collection_lib
        self._verbose = verbose
        self._abort = False
        self._done = False
        self._id = next(self._ids)
        self._run = None

    # thread_task interface (threads/thread_task.h)
    def id(self):
        return self._id

    def abort(self):
        self._abort = True

    def done(self):
        return self._done

    def _say(self, msg):
        if self._verbose:
            print(f"task #{self._id} wavefront_generator{msg}", flush=True)

    # the task body, in three parts so that wavefront_batch can interleave fans
    def _start(self):
        """wavefront_generator.cc:52-94: returns False if aborted before."""
        if self._abort:
            self._say(" *** aborted before execution ***")
            return False
        self._say(f": {self._source.description()} for {self._time_maximum} secs")
        if self._factory is None:
            from .waveq3d import wave_queue
            factory = wave_queue
        else:
            factory = self._factory
        wave = factory(self._ocean, self._frequencies, self._source_position, self._de_fan,
                       self._az_fan, self._time_step, self._target_positions)
        wave.intensity_threshold(self._intensity_threshold)
        wave.max_bottom(self._max_bottom)
        wave.max_surface(self._max_surface)
        rays = None
        if self._target_positions is not None:
            rays = eigenray_collection(self._frequencies, self._source_position,
                                       self._target_positions, self._source.keyID(),
                                       self._targetIDs, lib=self._collection_lib)
            if self._target_positions.size1() > 0 and self._target_positions.size2() > 0:
                wave.add_eigenray_listener(rays)
        verbs = eigenverb_collection(self._ocean.num_volume())
        if self._source.compute_reverb():
            wave.add_eigenverb_listener(verbs)
        if self._wavefront_file:
            self._say(f": writing wavefronts to {self._wavefront_file}")
            wave.init_netcdf(self._wavefront_file)
            wave.save_netcdf()
        self._run = {"wave": wave, "rays": rays, "verbs": verbs,
                     "left": _steps_to(self._time_step, self._time_maximum)}
        return True

    def _chunk(self, chunk):
        return 1 if self._wavefront_file else max(1, min(int(chunk), self._run["left"]))

    def _after(self, n):
        """After n steps completed: the abort check and the log record of
        wavefront_generator.cc:98-111.  Returns False when the task ended."""
        r = self._run
        r["left"] -= n
        wave = r["wave"]
        if self._abort:
            self._say(" *** aborted during execution ***")
            if self._wavefront_file:
                wave.close_netcdf()
            self._release()
            return False
        if self._wavefront_file:
            wave.save_netcdf()
        if r["left"] <= 0:
            self._finish()
            return False
        return True

    def _finish(self):
        """wavefront_generator.cc:113-126."""
        r = self._run
        if self._wavefront_file:
            r["wave"].close_netcdf()
        if r["rays"] is not None:
            r["rays"].sum_eigenrays()
        self._done = True
        self._release()
        self._source.notify_wavefront_listeners(self._source, r["rays"], r["verbs"])
        self._say(": done")

    def _release(self):
        wave = self._run["wave"]
        if hasattr(wave, "close"):
            wave.close()
        self._run["wave"] = None

    def run(self, chunk=1):
        """wavefront_generator.cc:52-127 on its own.  chunk > 1 queues that
        many steps at a time (abort is then checked per chunk)."""
        if not self._start():
            return
        while True:
            n = self._chunk(chunk)
            wave = self._run["wave"]
            wave.step_begin(n)
            wave.step_end(n)
            if not self._after(n):
                return


class wavefront_batch:
    """Runs many wavefront_generator tasks together on one GPU (the part the
    reference's thread_pool plays for CPU threads).

    Each round queues up to `chunk` steps on every live fan (step_begin: the
    launches return at once and the fans' streams run side by side), then
    completes them fan by fan (step_end: wait, replay listener calls) and
    checks each generator's abort flag.  Generators end independently."""

    def __init__(self, generators, chunk=16):
        self._gens = list(generators)
        self._chunk = int(chunk)

    def run(self):
        live = [g for g in self._gens if g._start()]
        while live:
            counts = []
            for g in live:
                n = g._chunk(self._chunk)
                g._run["wave"].step_begin(n)
                counts.append(n)
            still = []
            for g, n in zip(live, counts):
                g._run["wave"].step_end(n)
                if g._after(n):
                    still.append(g)
            live = still
