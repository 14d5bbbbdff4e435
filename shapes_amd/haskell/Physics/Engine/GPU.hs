{-# LANGUAGE ForeignFunctionInterface #-}
{-# LANGUAGE RecordWildCards          #-}
-- | The MI355X contact-physics step behind the engine API of
-- Physics.Engine.Main (shapes/src/Physics/Engine/Main.hs:32-98), over the C
-- ABI of include/shapes_gpu.h (libshapes_gpu.so).
--
-- The engine state is an opaque shp_batch handle instead of
-- (World, EngineCache, External).  A scene is marshalled once into the C
-- description: the data makeWorld, makeWorldObj, makePhysicalObj,
-- makeHull / makeRectangleHull / makeCircle and makeConstantAccel build
-- (shapes/src/Physics/Engine.hs:27-54, World.hs:118-149).  The solver is the
-- drop-in one (reference order, FP64, bit-exact).
module Physics.Engine.GPU
  ( ShapeDesc (..), BodyDesc (..), SceneDesc (..), GpuEngine
  , initEngine, changeScene, updateWorld, stepWorld, runWorld
  , readPhysObj, append, delete, filled
  , worldContacts, worldAabbs, dumpWorld
  , worldDescBytes, worldDescCapacityOffset, configBytes
  , bodyDescBytes, bodyDescVerticesOffset
  ) where

import           Control.Monad         (forM_, when, zipWithM_)
import qualified Data.ByteString       as BS
import           Data.IORef
import           Data.List             (zip6)
import           Foreign
import           Foreign.C.String
import           Foreign.C.Types

-- | A shape: a convex hull (local space, counter-clockwise, 3..8 vertices)
-- or a circle of the given radius (makeCircle, Engine.hs:53-54).
data ShapeDesc = HullDesc [(Double, Double)]
               | CircleDesc Double

-- | A body as makeWorldObj (makePhysicalObj vel rotvel pos rotpos (m, i))
-- mu bounce shape builds it (Engine.hs:27-39, World.hs:142-149).
data BodyDesc = BodyDesc
  { bVel :: (Double, Double), bRotVel :: Double, bPos :: (Double, Double)
  , bRotPos :: Double, bMass :: (Double, Double), bMu :: Double
  , bBounce :: Double, bShape :: ShapeDesc }

-- | Scene (Scenes/Scene.hs:13-17): the bodies in fromList order, the world's
-- slot capacity (World.new, World.hs:61-75; 0 for the body count), the
-- constant acceleration of makeConstantAccel, ContactBehavior, dt and the
-- number of improveWorld sweeps (2 in Main.hs:81-82).
data SceneDesc = SceneDesc
  { sBodies :: [BodyDesc], sCapacity :: Int, sAccel :: (Double, Double)
  , sBaumgarte :: Double, sSlop :: Double, sTimestep :: Double, sSweeps :: Int }

data ShpBatch

-- | EngineState: the handle and the world's slot count.
data GpuEngine = GpuEngine (IORef (ForeignPtr ShpBatch)) (IORef Int)

-- | shp_world_desc: int32 n_bodies (then 4 bytes of padding), 14 pointers
-- (px py theta vx vy omega mass inertia mu bounce hull_off hull_x hull_y
-- radius), int32 capacity, padded to 8 bytes.
worldDescBytes, worldDescCapacityOffset :: Int
worldDescBytes = 128
worldDescCapacityOffset = 120

-- | shp_config: dt baumgarte slop gravity_x gravity_y (double), then
-- iterations solver warm_start max_pairs_per_body broadphase solver_threads
-- (int32).
configBytes :: Int
configBytes = 64

-- | shp_body_desc: px py theta vx vy omega mass inertia mu bounce radius
-- (double), int32 n_vertices (then 4 bytes of padding), hull_x, hull_y.
bodyDescBytes, bodyDescVerticesOffset :: Int
bodyDescBytes = 112
bodyDescVerticesOffset = 88

foreign import ccall unsafe "shp_batch_create"
  c_create :: Ptr () -> CInt -> Ptr () -> Ptr (Ptr ShpBatch) -> IO CInt
foreign import ccall unsafe "&shp_batch_destroy"
  c_destroy :: FunPtr (Ptr ShpBatch -> IO ())
foreign import ccall safe "shp_batch_step"
  c_step :: Ptr ShpBatch -> CInt -> IO CInt
foreign import ccall safe "shp_batch_read_bodies"
  c_read_bodies :: Ptr ShpBatch -> CInt -> Ptr CDouble -> Ptr CDouble -> Ptr CDouble
                -> Ptr CDouble -> Ptr CDouble -> Ptr CDouble -> IO CInt
foreign import ccall safe "shp_batch_append"
  c_append :: Ptr ShpBatch -> CInt -> Ptr () -> Ptr CInt -> IO CInt
foreign import ccall safe "shp_batch_delete"
  c_delete :: Ptr ShpBatch -> CInt -> CInt -> IO CInt
foreign import ccall unsafe "shp_batch_filled"
  c_filled :: Ptr ShpBatch -> CInt -> Ptr CInt -> IO CInt
foreign import ccall safe "shp_batch_probe_contacts"
  c_probe :: Ptr ShpBatch -> IO CInt
foreign import ccall safe "shp_batch_read_contacts"
  c_read_contacts :: Ptr ShpBatch -> CInt -> Ptr CInt -> Ptr CInt -> Ptr CInt
                  -> Ptr CDouble -> Ptr CDouble -> Ptr CInt -> IO CInt
foreign import ccall safe "shp_batch_read_aabbs"
  c_read_aabbs :: Ptr ShpBatch -> CInt -> Ptr CInt -> Ptr CInt -> Ptr CDouble -> IO CInt
foreign import ccall safe "shp_batch_dump"
  c_dump :: Ptr ShpBatch -> CInt -> CUInt -> Ptr Word8 -> Int64 -> Ptr Int64 -> IO CInt
foreign import ccall unsafe "shp_last_error"
  c_last_error :: IO CString

check :: CInt -> IO ()
check 0 = return ()
check s = do
  msg <- c_last_error >>= peekCString
  ioError (userError ("shapes_gpu status " ++ show s ++ ": " ++ msg))

withDescs :: SceneDesc -> (Ptr () -> Ptr () -> IO a) -> IO a
withDescs SceneDesc{..} act = do
  let bs = sBodies
      col f = map (realToFrac . f) bs :: [CDouble]
      verts b = case bShape b of HullDesc vs -> vs; CircleDesc _ -> []
      radius b = case bShape b of HullDesc _ -> 0; CircleDesc r -> r
      counts = map (length . verts) bs
      offs = map fromIntegral (scanl (+) 0 counts) :: [CInt]
      hx = [realToFrac x | b <- bs, (x, _) <- verts b] :: [CDouble]
      hy = [realToFrac y | b <- bs, (_, y) <- verts b] :: [CDouble]
      cols = [ col (fst . bPos), col (snd . bPos), col bRotPos, col (fst . bVel)
             , col (snd . bVel), col bRotVel, col (fst . bMass), col (snd . bMass)
             , col bMu, col bBounce ]
  withArrays cols $ \ps ->
    withArray offs $ \po -> withArray hx $ \px -> withArray hy $ \py ->
    withArray (map (realToFrac . radius) bs :: [CDouble]) $ \pr ->
    allocaBytes worldDescBytes $ \wd -> allocaBytes configBytes $ \cfg -> do
      pokeByteOff wd 0 (fromIntegral (length bs) :: Int32)
      forM_ (zip [0 ..] (ps ++ [castPtr po, castPtr px, castPtr py, castPtr pr])) $ \(k, q) ->
        pokeByteOff wd (8 + 8 * k) (q :: Ptr ())
      pokeByteOff wd worldDescCapacityOffset (fromIntegral sCapacity :: Int32)
      forM_ (zip [0 ..] [sTimestep, sBaumgarte, sSlop, fst sAccel, snd sAccel]) $ \(k, x) ->
        pokeByteOff cfg (8 * k) (realToFrac x :: CDouble)
      -- sweeps; the ordered FP64 solver (bit-exact drop-in); zero warm
      -- start; default capacity, broadphase and lanes
      forM_ (zip [0 ..] [sSweeps, 0, 0, 0, 0, 0]) $ \(k, x) ->
        pokeByteOff cfg (40 + 4 * k) (fromIntegral x :: Int32)
      act wd cfg
  where
    withArrays [] f = f []
    withArrays (a : as) f = withArray a $ \p -> withArrays as (f . (castPtr p :))

create :: SceneDesc -> IO (ForeignPtr ShpBatch)
create sc = withDescs sc $ \wd cfg -> alloca $ \ph -> do
  check =<< c_create wd 1 cfg ph
  peek ph >>= newForeignPtr c_destroy

slots :: SceneDesc -> Int
slots sc = max (sCapacity sc) (length (sBodies sc))

withEngine :: GpuEngine -> (Ptr ShpBatch -> IO a) -> IO a
withEngine (GpuEngine h _) act = readIORef h >>= \fp -> withForeignPtr fp act

-- | initEngine (Main.hs:43-46): the world, an empty cache, the externals.
initEngine :: SceneDesc -> IO GpuEngine
initEngine sc = GpuEngine <$> (create sc >>= newIORef) <*> newIORef (slots sc)

-- | changeScene (Main.hs:48-51): replace the state by a new scene's.
changeScene :: SceneDesc -> GpuEngine -> IO ()
changeScene sc (GpuEngine h n) = do
  create sc >>= writeIORef h
  writeIORef n (slots sc)

-- | updateWorld (Main.hs:71-86) and stepWorld n (:88-89).
updateWorld :: GpuEngine -> IO ()
updateWorld = stepWorld 1

stepWorld :: Int -> GpuEngine -> IO ()
stepWorld steps e = when (steps > 0) $
  withEngine e $ \p -> check =<< c_step p (fromIntegral steps)

-- | runWorld dt scene n (Main.hs:97-98); returns the final state.
runWorld :: Double -> SceneDesc -> Int -> IO GpuEngine
runWorld dt sc steps = do
  e <- initEngine sc { sTimestep = dt }
  stepWorld steps e
  return e

-- | readPhysObj (World.hs:93-94) of every slot: (pos, rotPos, vel, rotVel).
-- An empty slot keeps the state it had when it was emptied.
readPhysObj :: GpuEngine -> IO [((Double, Double), Double, (Double, Double), Double)]
readPhysObj e@(GpuEngine _ nref) = do
  n <- readIORef nref
  withEngine e $ \p ->
    allocaArray n $ \px -> allocaArray n $ \py -> allocaArray n $ \th ->
    allocaArray n $ \vx -> allocaArray n $ \vy -> allocaArray n $ \w -> do
      check =<< c_read_bodies p 0 px py th vx vy w
      [a, b, c, d, f, g] <- mapM (fmap (map realToFrac) . peekArray n) [px, py, th, vx, vy, w]
      return [((x, y), t, (u, v), o) | (x, y, t, u, v, o) <- zip6 a b c d f g]

-- | World.append (World.hs:77-84): the body into the first empty slot of
-- the free list; returns the slot.
append :: GpuEngine -> BodyDesc -> IO Int
append e BodyDesc{..} = do
  let (vs, r) = case bShape of
        HullDesc v -> (v, 0)
        CircleDesc c -> ([], c)
  withEngine e $ \p ->
    withArray ([realToFrac x | (x, _) <- vs] :: [CDouble]) $ \hx ->
    withArray ([realToFrac y | (_, y) <- vs] :: [CDouble]) $ \hy ->
    allocaBytes bodyDescBytes $ \bd -> alloca $ \slot -> do
      zipWithM_ (\k x -> pokeByteOff bd (8 * k) (realToFrac x :: CDouble)) [0 ..]
        [ fst bPos, snd bPos, bRotPos, fst bVel, snd bVel, bRotVel
        , fst bMass, snd bMass, bMu, bBounce, r ]
      pokeByteOff bd bodyDescVerticesOffset (fromIntegral (length vs) :: Int32)
      pokeByteOff bd (bodyDescVerticesOffset + 8) hx
      pokeByteOff bd (bodyDescVerticesOffset + 16) hy
      check =<< c_append p 0 bd slot
      fromIntegral <$> peek slot

-- | World.delete (World.hs:86-87, EmptiesVector.safeDelete): fails on an
-- empty slot.
delete :: GpuEngine -> Int -> IO ()
delete e i = withEngine e $ \p -> check =<< c_delete p 0 (fromIntegral i)

-- | World.filled (World.hs:90-91).
filled :: GpuEngine -> IO Int
filled e = withEngine e $ \p -> alloca $ \n -> do
  check =<< c_filled p 0 n
  fromIntegral <$> peek n

chunks :: Int -> [a] -> [[a]]
chunks k = takeWhile (not . null) . map (take k) . iterate (drop k)

-- | The demo's worldContacts (shapes-demo IOWorld.hs:74-81): the contacts
-- culledKeys and prepareFrame find in the world as it stands, as
-- (normal, centre, depth) (Contact/Types.hs:28-34).
worldContacts :: GpuEngine -> IO [((Double, Double), (Double, Double), Double)]
worldContacts e = withEngine e $ \p -> do
  check =<< c_probe p
  n <- alloca $ \pn -> do
    check =<< c_read_contacts p 0 pn nullPtr nullPtr nullPtr nullPtr nullPtr
    fromIntegral <$> peek pn
  allocaArray (5 * n) $ \g -> alloca $ \pn -> do
    check =<< c_read_contacts p 0 pn nullPtr nullPtr g nullPtr nullPtr
    xs <- map realToFrac <$> peekArray (5 * n) g
    return [((a, b), (c, d), f) | [a, b, c, d, f] <- chunks 5 xs]

-- | The demo's worldAabbs (IOWorld.hs:86-90): per filled slot
-- ((minx, maxx), (miny, maxy)).
worldAabbs :: GpuEngine -> IO [((Double, Double), (Double, Double))]
worldAabbs e = withEngine e $ \p -> do
  n <- alloca $ \pn -> do
    check =<< c_read_aabbs p 0 pn nullPtr nullPtr
    fromIntegral <$> peek pn
  allocaArray (4 * n) $ \bx -> allocaArray n $ \sl -> alloca $ \pn -> do
    check =<< c_read_aabbs p 0 pn sl bx
    xs <- map realToFrac <$> peekArray (4 * n) bx
    return [((a, b), (c, d)) | [a, b, c, d] <- chunks 4 xs]

-- | shp_batch_dump of every section (bodies, AABBs, pairs, contacts,
-- cache): the byte format the golden fixtures use.
dumpWorld :: GpuEngine -> IO BS.ByteString
dumpWorld e = withEngine e $ \p -> alloca $ \psz -> do
  check =<< c_dump p 0 31 nullPtr 0 psz
  sz <- peek psz
  allocaBytes (fromIntegral sz) $ \buf -> do
    check =<< c_dump p 0 31 buf sz psz
    BS.packCStringLen (castPtr buf, fromIntegral sz)
